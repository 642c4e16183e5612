// CISD trial with restricted walkers (trot/trial/cisd.py, trot/meas/cisd.py), FP64.
//
// The reference determinant occupies the first n = nocc_full orbitals, so the half Green's function
// is GT = w w[:n]^-1 (N x n, the RHF kernels with C = I[:, :n]).  Active blocks: occupied rows
// core .. core+na-1, virtual rows n .. n+va-1 (ov = na va amplitudes).  With
//   x[a,t]  = GT[n+t, core+a]                       (green_occ of trot/meas/cisd.py:24-44)
//   D[pt,qu] = 2 ci2[p,t,q,u] - ci2[p,u,q,t]        (both doubles contractions at once)
//   y = x D,  c1 = <ci1, x>,  c2 = <y, x>,  Omega = 1 + 2 c1 + c2,  Bs = (ci1 + y) / Omega
//   Ps[i,a] = sum_t GT[n+t, i] Bs[a,t] (i < n),  -Bs[a, i-n] (n <= i < n+va),  0 otherwise
//   gamma  = G0 - Ps G_act,  G0[i,p] = GT[i,p] (p < n),  G_act[a,j] = GT[j, core+a]
// the reference's quantities become
//   overlap = det(w[:n])^2 Omega                                   (trial/cisd.py:148-166)
//   fb_g    = 2 sum_ij L_g[i,j] gamma[i,j]                          (meas/cisd.py:548-564)
//   rdm1    = gamma                                                 (meas/cisd.py:772-820)
// and, with R_g = L_g GT (N x n) for every Cholesky vector and R_h = h1 GT,
//   glgp_g[p,t] = sum_{i<n} R_g[i,p] GT[n+t,i] - R_g[n+t,p]         (gl greenp of :745-755)
//   E = h0 + 2 tr R_h - 2 <Ps, R_h[:,act]> + sum_g (2 (tr R_g)^2 - tr(R_g[:n] R_g[:n]))
//       - 4 sum_g tr R_g <Ps, R_g[:,act]> + 2 sum_g <glgp_g, R_g[:n,act] Bs>
//       + sum_g glgp_g[act]^T D glgp_g[act] / Omega                  (meas/cisd.py:608-769)
// which is the reference's sum of terms regrouped; it uses that L_g and h1 are symmetric, as they are
// for every Hamiltonian the reference builds (real orbitals).
//
//   cisd_dmat_kernel      D from ci2
//   pack_sym_kernel       upper triangles of symmetric matrices, row-major (the force-bias K index)
//   cisd_gather_kernel    x from GT
//   cisd_overlap_kernel   c1, c2, Omega, overlap, Bs (one workgroup per walker)
//   cisd_ps_top_kernel    Ps rows < nocc; cisd_ps_rest_kernel the rest
//   cisd_gamma_tile_kernel gamma, full or as packed symmetric sums gamma[i,j] + gamma[j,i]
//   cisd_permute_kernel   GT of a walker chunk side by side (N x Wc n) for one R GEMM
//   cisd_terms_kernel     per (walker, g): tr R_g, tr(R_g R_g), <glgp_g, H_g>, <Ps, R_g[:,act]>, x_g
//   cisd_rowdot_kernel    x_g^T (x_g D)
//   cisd_combine_kernel   the local energy of each walker (fixed summation order)
#include "capi_util.hpp"
#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace {

__global__ __launch_bounds__(256) void cisd_dmat_kernel(const double* __restrict__ ci2, int na, int va,
                                                       double* __restrict__ D) {
  const long ov = (long)na * va;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= ov * ov) return;
  const long r = e / ov, c = e - r * ov;
  const int p = (int)(r / va), t = (int)(r - (long)p * va);
  const int q = (int)(c / va), u = (int)(c - (long)q * va);
  // ci2[p,t,q,u] and ci2[p,u,q,t]
  const double direct = ci2[(((long)p * va + t) * na + q) * va + u];
  const double exch = ci2[(((long)p * va + u) * na + q) * va + t];
  D[e] = 2.0 * direct - exch;
}

__global__ __launch_bounds__(256) void pack_sym_kernel(const double* __restrict__ m, int N, long np,
                                                      double* __restrict__ out) {
  const long g = blockIdx.y;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)N * N) return;
  const int i = (int)(e / N), j = (int)(e - (long)i * N);
  if (j < i) return;
  const long idx = (long)i * N - ((long)i * (i - 1)) / 2 + (j - i);
  out[g * np + idx] = m[g * (long)N * N + e];
}

__global__ __launch_bounds__(256) void cisd_gather_kernel(const double* __restrict__ gt, int N, int n, int core,
                                                         int na, int va, double* __restrict__ x) {
  const long w = blockIdx.y;
  const long ov = (long)na * va;
  const long k = (long)blockIdx.x * 256 + threadIdx.x;
  if (k >= ov) return;
  const int a = (int)(k / va), t = (int)(k - (long)a * va);
  cstore(x, w * ov + k, cload(gt, w * N * n + (long)(n + t) * n + core + a));
}

// One workgroup per walker.  det: complex det(w[:n]) of one spin; ovl (optional) = det^2 Omega.
__global__ __launch_bounds__(256) void cisd_overlap_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                          const double* __restrict__ ci1, const double* __restrict__ det,
                                                          long ov, double* __restrict__ omega,
                                                          double* __restrict__ ovl, double* __restrict__ bs) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const long w = blockIdx.x;
  double c1r = 0.0, c1i = 0.0, c2r = 0.0, c2i = 0.0;
  for (long k = tid; k < ov; k += 256) {
    const cplx xv = cload(x, w * ov + k), yv = cload(y, w * ov + k);
    const double c = ci1[k];
    c1r += c * xv.re;
    c1i += c * xv.im;
    const cplx p = cmul(yv, xv);
    c2r += p.re;
    c2i += p.im;
  }
  c1r = block_sum<256>(c1r, red);
  c1i = block_sum<256>(c1i, red);
  c2r = block_sum<256>(c2r, red);
  c2i = block_sum<256>(c2i, red);
  const cplx om = {1.0 + 2.0 * c1r + c2r, 2.0 * c1i + c2i};
  if (tid == 0) {
    cstore(omega, w, om);
    if (ovl) {
      const cplx d = cload(det, w);
      cstore(ovl, w, cmul(cmul(d, d), om));
    }
  }
  const cplx inv = cinv(om);
  for (long k = tid; k < ov; k += 256) {
    const cplx yv = cload(y, w * ov + k);
    cstore(bs, w * ov + k, cmul(cplx{ci1[k] + yv.re, yv.im}, inv));
  }
}

// Ps rows i < n: Ps[i,a] = sum_t GT[n+t, i] Bs[a,t].  One workgroup per walker; t in chunks of PS_TT
// staged in LDS (GT rows n+t: PS_TT x n; Bs columns: na x PS_TT); (i, a) pairs in registers (n na <= 4096).
constexpr int PS_TT = 32;
__global__ __launch_bounds__(256) void cisd_ps_top_kernel(const double* __restrict__ gt, const double* __restrict__ bs,
                                                         int N, int n, int na, int va, double* __restrict__ ps) {
  extern __shared__ double sh[];
  double* gv = sh;                                // [PS_TT][n + 1] complex
  double* bc = sh + 2 * PS_TT * (n + 1);          // [na][PS_TT + 1] complex
  const int tid = threadIdx.x;
  const long w = blockIdx.x;
  const long ov = (long)na * va;
  const double* g = gt + 2 * w * N * n;
  const double* b = bs + 2 * w * ov;
  const int npair = n * na;
  cplx acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = {0.0, 0.0};
  for (int t0 = 0; t0 < va; t0 += PS_TT) {
    const int tt = va - t0 < PS_TT ? va - t0 : PS_TT;
    for (int e = tid; e < tt * n; e += 256) {
      const int t = e / n, i = e - t * n;
      cstore(gv, (long)t * (n + 1) + i, cload(g, (long)(n + t0 + t) * n + i));
    }
    for (int e = tid; e < na * tt; e += 256) {
      const int a = e / tt, t = e - a * tt;
      cstore(bc, (long)a * (PS_TT + 1) + t, cload(b, (long)a * va + t0 + t));
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int e = tid + k * 256;
      if (e < npair) {
        const int i = e / na, a = e - i * na;
        cplx s = acc[k];
        for (int t = 0; t < tt; ++t)
          s = cadd(s, cmul(cload(gv, (long)t * (n + 1) + i), cload(bc, (long)a * (PS_TT + 1) + t)));
        acc[k] = s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int e = tid + k * 256;
    if (e < npair) cstore(ps, w * N * na + e, acc[k]);
  }
}

// Ps rows i >= n: -Bs[a, i - n] for i < n + va, zero beyond.
__global__ __launch_bounds__(256) void cisd_ps_rest_kernel(const double* __restrict__ bs, int N, int n, int na, int va,
                                                          double* __restrict__ ps) {
  const long w = blockIdx.y;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)(N - n) * na) return;
  const int i = n + (int)(e / na), a = (int)(e % na);
  cplx v = {0.0, 0.0};
  if (i < n + va) {
    const cplx x = cload(bs, (w * na + a) * va + (i - n));
    v = {-x.re, -x.im};
  }
  cstore(ps, w * N * na + (long)n * na + e, v);
}

// gamma = G0 - Ps G_act on (TS x TS) tiles (bi <= bj) of one walker: the tile and its mirror from four
// LDS panels (Ps rows and GT active columns of both row blocks).  packed: gamma[i,j] + gamma[j,i] (i < j) and
// gamma[i,i] at the trot_pack_sym index; else both gamma[i,j] and gamma[j,i] into the full matrix.
template <int TS>
__global__ __launch_bounds__(256) void cisd_gamma_tile_kernel(const double* __restrict__ gt,
                                                             const double* __restrict__ ps, int N, int n, int core,
                                                             int na, int nt, int packed, double* __restrict__ out) {
  extern __shared__ double sh[];
  const int ld = na + 1;
  double* Pi = sh;
  double* Pj = Pi + 2 * TS * ld;
  double* Ai = Pj + 2 * TS * ld;
  double* Aj = Ai + 2 * TS * ld;
  const int tid = threadIdx.x;
  const long w = blockIdx.y;
  int bi = 0, rem = blockIdx.x;
  while (rem >= nt - bi) {
    rem -= nt - bi;
    ++bi;
  }
  const int bj = bi + rem;
  const double* g = gt + 2 * w * N * n;
  const double* p = ps + 2 * w * N * na;
  for (int e = tid; e < TS * na; e += 256) {
    const int r = e / na, a = e - r * na;
    const int i = bi * TS + r, j = bj * TS + r;
    const cplx z = {0.0, 0.0};
    cstore(Pi, (long)r * ld + a, i < N ? cload(p, (long)i * na + a) : z);
    cstore(Pj, (long)r * ld + a, j < N ? cload(p, (long)j * na + a) : z);
    cstore(Ai, (long)r * ld + a, i < N ? cload(g, (long)i * n + core + a) : z);
    cstore(Aj, (long)r * ld + a, j < N ? cload(g, (long)j * n + core + a) : z);
  }
  __syncthreads();
  const long np = (long)N * (N + 1) / 2;
  for (int e = tid; e < TS * TS; e += 256) {
    const int ri = e / TS, rj = e - ri * TS;
    const int i = bi * TS + ri, j = bj * TS + rj;
    if (i >= N || j >= N || (bi == bj && rj < ri)) continue;
    cplx gij = j < n ? cload(g, (long)i * n + j) : cplx{0.0, 0.0};
    cplx gji = i < n ? cload(g, (long)j * n + i) : cplx{0.0, 0.0};
    for (int a = 0; a < na; ++a) {
      gij = csub(gij, cmul(cload(Pi, (long)ri * ld + a), cload(Aj, (long)rj * ld + a)));
      gji = csub(gji, cmul(cload(Pj, (long)rj * ld + a), cload(Ai, (long)ri * ld + a)));
    }
    if (packed) {
      const cplx v = i == j ? gij : cadd(gij, gji);
      cstore(out, w * np + (long)i * N - ((long)i * (i - 1)) / 2 + (j - i), v);
    } else {
      cstore(out, (w * N + i) * N + j, gij);
      cstore(out, (w * N + j) * N + i, gji);
    }
  }
}

// gtall[j, wl n + p] = gt[w0 + wl, j, p]
__global__ __launch_bounds__(256) void cisd_permute_kernel(const double* __restrict__ gt, int N, int n, int wc,
                                                          double* __restrict__ gtall) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  const long cols = (long)wc * n;
  if (e >= (long)N * cols) return;
  const long j = e / cols, c = e - j * cols;
  const long wl = c / n, p = c - wl * n;
  cstore(gtall, e, cload(gt, (wl * N + j) * n + p));
}

// One workgroup per (g, walker of the chunk).  R: ((G+1) N) x (wc n), row (g N + i), column (wl n + p); row
// block G holds h1 GT.  terms[(wl (G+1) + g) * 4 + k]: tr R_g, tr(R_g[:n] R_g[:n]), <glgp, H>, <Ps, R_g[:,act]>,
// with H = R_g[:n, act] Bs.  x_g (g < G) goes to xg[(wl G + g) ov + a va + t].  LDS: f = R_g[:n] (n x n) for the
// whole workgroup; GT rows n+t, R_g rows n+t and Bs columns t in chunks of tt.
__global__ __launch_bounds__(256) void cisd_terms_kernel(const double* __restrict__ R, const double* __restrict__ gt,
                                                        const double* __restrict__ bs, const double* __restrict__ ps,
                                                        int N, int n, int G, int core, int na, int va, int wc, int TT,
                                                        double* __restrict__ terms, double* __restrict__ xg) {
  extern __shared__ double sh[];
  __shared__ double red[256];
  const int ldn = n + 1;
  double* f = sh;                           // [n][n + 1]
  double* rv = f + 2 * n * ldn;             // [TT][n + 1]
  double* gv = rv + 2 * TT * ldn;           // [TT][n + 1]
  double* bc = gv + 2 * TT * ldn;           // [na][TT + 1]
  const int tid = threadIdx.x;
  const int g = blockIdx.x;
  const long wl = blockIdx.y;
  const long ldc = (long)wc * n;
  const double* Rg = R + 2 * ((long)g * N * ldc + wl * n);  // R_g[i,p] at (i ldc + p)
  const double* gw = gt + 2 * wl * N * n;
  const long ov = (long)na * va;
  const double* b = bs + 2 * wl * ov;
  const double* pw = ps + 2 * wl * N * na;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int e = tid; e < n * n; e += 256) {
    const int p = e / n, q = e - p * n;
    cstore(f, (long)p * ldn + q, cload(Rg, (long)p * ldc + q));
  }
  // <Ps, R_g[:, act]> straight from memory (each element used once)
  for (long e = tid; e < (long)N * na; e += 256) {
    const int i = (int)(e / na), a = (int)(e - (long)i * na);
    const cplx m = cmul(cload(pw, e), cload(Rg, (long)i * ldc + core + a));
    v[6] += m.re;
    v[7] += m.im;
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += 256) {
    const int p = e / n, q = e - p * n;
    const cplx rpq = cload(f, (long)p * ldn + q);
    if (p == q) {
      v[0] += rpq.re;
      v[1] += rpq.im;
    }
    const cplx m = cmul(rpq, cload(f, (long)q * ldn + p));
    v[2] += m.re;
    v[3] += m.im;
  }
  for (int t0 = 0; t0 < va; t0 += TT) {
    const int tt = va - t0 < TT ? va - t0 : TT;
    for (int e = tid; e < tt * n; e += 256) {
      const int t = e / n, i = e - t * n;
      cstore(rv, (long)t * ldn + i, cload(Rg, (long)(n + t0 + t) * ldc + i));
      cstore(gv, (long)t * ldn + i, cload(gw, (long)(n + t0 + t) * n + i));
    }
    for (int e = tid; e < na * tt; e += 256) {
      const int a = e / tt, t = e - a * tt;
      cstore(bc, (long)a * (TT + 1) + t, cload(b, (long)a * va + t0 + t));
    }
    __syncthreads();
    for (int e = tid; e < n * tt; e += 256) {
      const int p = e / tt, t = e - p * tt;
      cplx gl = cload(rv, (long)t * ldn + p);
      gl = {-gl.re, -gl.im};
      for (int i = 0; i < n; ++i) gl = cadd(gl, cmul(cload(f, (long)i * ldn + p), cload(gv, (long)t * ldn + i)));
      cplx h = {0.0, 0.0};
      for (int a = 0; a < na; ++a) h = cadd(h, cmul(cload(f, (long)p * ldn + core + a), cload(bc, (long)a * (TT + 1) + t)));
      const cplx m = cmul(gl, h);
      v[4] += m.re;
      v[5] += m.im;
      if (g < G && p >= core && p < core + na) cstore(xg, (wl * G + g) * ov + (long)(p - core) * va + t0 + t, gl);
    }
    __syncthreads();
  }
  double s[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = block_sum<256>(v[k], red);
  if (tid < 4) {
    const long o = (wl * (G + 1) + g) * 4 + tid;
    cstore(terms, o, cplx{s[2 * tid], s[2 * tid + 1]});
  }
}

// q[r] = sum_c Z[r,c] X[r,c]
__global__ __launch_bounds__(256) void cisd_rowdot_kernel(const double* __restrict__ Z, const double* __restrict__ X,
                                                         long ov, double* __restrict__ q) {
  __shared__ double red[256];
  const long r = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (long c = threadIdx.x; c < ov; c += 256) {
    const cplx m = cmul(cload(Z, r * ov + c), cload(X, r * ov + c));
    re += m.re;
    im += m.im;
  }
  re = block_sum<256>(re, red);
  im = block_sum<256>(im, red);
  if (threadIdx.x == 0) cstore(q, r, cplx{re, im});
}

// One workgroup per walker of the chunk; the sums over g run in a fixed order.
__global__ __launch_bounds__(256) void cisd_combine_kernel(const double* __restrict__ terms,
                                                          const double* __restrict__ q,
                                                          const double* __restrict__ omega, int G, double h0,
                                                          double* __restrict__ e) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const long wl = blockIdx.x;
  // s_a = sum_g (2 lg^2 - exch - 4 lg lmc + 2 T),  s_q = sum_g q
  double ar = 0.0, ai = 0.0, qr = 0.0, qi = 0.0;
  for (int g = tid; g < G; g += 256) {
    const long o = (wl * (G + 1) + g) * 4;
    const cplx lg = cload(terms, o), ex = cload(terms, o + 1), tt = cload(terms, o + 2), lm = cload(terms, o + 3);
    const cplx l2 = cmul(lg, lg), llm = cmul(lg, lm);
    ar += 2.0 * l2.re - ex.re - 4.0 * llm.re + 2.0 * tt.re;
    ai += 2.0 * l2.im - ex.im - 4.0 * llm.im + 2.0 * tt.im;
    const cplx qv = cload(q, wl * G + g);
    qr += qv.re;
    qi += qv.im;
  }
  ar = block_sum<256>(ar, red);
  ai = block_sum<256>(ai, red);
  qr = block_sum<256>(qr, red);
  qi = block_sum<256>(qi, red);
  if (tid == 0) {
    const long o = (wl * (G + 1) + G) * 4;
    const cplx hg = cload(terms, o), hm = cload(terms, o + 3);
    const cplx qo = cmul(cplx{qr, qi}, cinv(cload(omega, wl)));
    cstore(e, wl, cplx{h0 + 2.0 * hg.re - 2.0 * hm.re + ar + qo.re, 2.0 * hg.im - 2.0 * hm.im + ai + qo.im});
  }
}

inline unsigned blocks_of(long n) { return (unsigned)((n + 255) / 256); }

hipError_t set_lds(const void* f, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

int check_shape(int W, int N, int n, int core, int na, int va) {
  if (W < 1 || n < 1 || na < 1 || va < 1 || core < 0 || core + na != n || n + va > N)
    return fail(TROT_ERR_INVALID, "cisd: bad sizes (need core + na = nocc, nocc + nvir <= norb, na, nvir >= 1)");
  if (n > TROT_MAX_NOCC) return fail(TROT_ERR_UNSUPPORTED, "cisd: nocc > 64");
  if ((long)na * va > (1L << 30)) return fail(TROT_ERR_UNSUPPORTED, "cisd: na nvir too large");
  return TROT_OK;
}

// y (rows x ov) = x (rows x ov) D (ov x ov)
hipError_t times_d(const double* x, const double* D, double* y, long rows, long ov, hipStream_t s) {
  GemmP p{};
  p.A = x; p.lda = ov; p.sA = 0;
  p.B = D; p.ldb = ov; p.sB = 0;
  p.C = y; p.ldc = ov; p.sC = 0;
  p.M = (int)rows; p.N = (int)ov; p.K = (int)ov; p.ar = 1.0;
  return launch_gemm_f64<64, 64, true, false, true, true, false, EPI_STORE>(p, 1, s);
}

}  // namespace

extern "C" {

int trot_cisd_doubles_matrix(const double* ci2, int na, int va, double* D, void* stream) {
  if (!ci2 || !D || na < 1 || va < 1) return fail(TROT_ERR_INVALID, "cisd_doubles_matrix: bad arguments");
  const long ov = (long)na * va;
  if (ov > (1L << 20)) return fail(TROT_ERR_UNSUPPORTED, "cisd_doubles_matrix: na nvir > 2^20");
  hipLaunchKernelGGL(cisd_dmat_kernel, dim3(blocks_of(ov * ov)), dim3(256), 0, (hipStream_t)stream, ci2, na, va, D);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_pack_sym(const double* m, int count, int N, double* out, void* stream) {
  if (!m || !out || count < 1 || N < 1) return fail(TROT_ERR_INVALID, "pack_sym: bad arguments");
  if (count > 65535) return fail(TROT_ERR_UNSUPPORTED, "pack_sym: more than 65535 matrices");
  const long np = (long)N * (N + 1) / 2;
  hipLaunchKernelGGL(pack_sym_kernel, dim3(blocks_of((long)N * N), count), dim3(256), 0, (hipStream_t)stream, m, N, np,
                     out);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

size_t trot_cisd_overlap_workspace(int W, int na, int va) {
  if (W < 1 || na < 1 || va < 1) return 0;
  return 2 * align256((size_t)W * na * va * 16);
}

int trot_cisd_overlap(const double* gt, const double* det, const double* ci1, const double* D, int W, int N, int n,
                      int core, int na, int va, double* omega, double* ovl, double* bs, void* ws, size_t ws_bytes,
                      void* stream) {
  if (!gt || !ci1 || !D || !omega || !bs || (ovl && !det)) return fail(TROT_ERR_INVALID, "cisd_overlap: null pointer");
  if (int rc = check_shape(W, N, n, core, na, va)) return rc;
  if (W > 65535) return fail(TROT_ERR_UNSUPPORTED, "cisd_overlap: more than 65535 walkers per call");
  if (!ws || ws_bytes < trot_cisd_overlap_workspace(W, na, va)) return fail(TROT_ERR_WORKSPACE, "cisd_overlap: workspace");
  const long ov = (long)na * va;
  Carve cv{(char*)ws, ws_bytes, 0};
  double* x = cv.take<double>((size_t)W * ov * 2);
  double* y = cv.take<double>((size_t)W * ov * 2);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(cisd_gather_kernel, dim3(blocks_of(ov), W), dim3(256), 0, s, gt, N, n, core, na, va, x);
  HIPCHK(hipGetLastError());
  HIPCHK(times_d(x, D, y, W, ov, s));
  hipLaunchKernelGGL(cisd_overlap_kernel, dim3(W), dim3(256), 0, s, x, y, ci1, det, ov, omega, ovl, bs);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_cisd_ps(const double* gt, const double* bs, int W, int N, int n, int core, int na, int va, double* ps,
                 void* stream) {
  if (!gt || !bs || !ps) return fail(TROT_ERR_INVALID, "cisd_ps: null pointer");
  if (int rc = check_shape(W, N, n, core, na, va)) return rc;
  if (W > 65535) return fail(TROT_ERR_UNSUPPORTED, "cisd_ps: more than 65535 walkers per call");
  hipStream_t s = (hipStream_t)stream;
  const size_t lds = (size_t)(PS_TT * (n + 1) + na * (PS_TT + 1)) * 16;
  HIPCHK(set_lds((const void*)cisd_ps_top_kernel, lds));
  hipLaunchKernelGGL(cisd_ps_top_kernel, dim3(W), dim3(256), lds, s, gt, bs, N, n, na, va, ps);
  HIPCHK(hipGetLastError());
  if (N > n) {
    hipLaunchKernelGGL(cisd_ps_rest_kernel, dim3(blocks_of((long)(N - n) * na), W), dim3(256), 0, s, bs, N, n, na, va,
                       ps);
    HIPCHK(hipGetLastError());
  }
  return TROT_OK;
}

int trot_cisd_gamma(const double* gt, const double* ps, int W, int N, int n, int core, int na, int packed, double* out,
                    void* stream) {
  if (!gt || !ps || !out) return fail(TROT_ERR_INVALID, "cisd_gamma: null pointer");
  if (W < 1 || n < 1 || na < 1 || core < 0 || core + na != n || n > N)
    return fail(TROT_ERR_INVALID, "cisd_gamma: bad sizes");
  if (W > 65535) return fail(TROT_ERR_UNSUPPORTED, "cisd_gamma: more than 65535 walkers per call");
  if (na <= 32) {
    const int nt = (N + 31) / 32;
    const size_t lds = (size_t)4 * 32 * (na + 1) * 16;
    HIPCHK(set_lds((const void*)cisd_gamma_tile_kernel<32>, lds));
    hipLaunchKernelGGL(cisd_gamma_tile_kernel<32>, dim3(nt * (nt + 1) / 2, W), dim3(256), lds, (hipStream_t)stream, gt,
                       ps, N, n, core, na, nt, packed, out);
  } else {
    const int nt = (N + 15) / 16;
    const size_t lds = (size_t)4 * 16 * (na + 1) * 16;
    HIPCHK(set_lds((const void*)cisd_gamma_tile_kernel<16>, lds));
    hipLaunchKernelGGL(cisd_gamma_tile_kernel<16>, dim3(nt * (nt + 1) / 2, W), dim3(256), lds, (hipStream_t)stream, gt,
                       ps, N, n, core, na, nt, packed, out);
  }
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

size_t trot_cisd_energy_workspace(int wc, int N, int n, int G, int na, int va, int b3) {
  if (wc < 1 || N < 1 || n < 1 || G < 1 || na < 1 || va < 1) return 0;
  const size_t ov = (size_t)na * va;
  size_t bytes = align256((size_t)N * wc * n * 16) + align256((size_t)(G + 1) * N * wc * n * 16) +
                 2 * align256((size_t)wc * G * ov * 16) + align256((size_t)wc * (G + 1) * 4 * 16) +
                 align256((size_t)wc * G * 16);
  if (b3) bytes += align256(trot_b3_planes_bytes(wc * G, (int)ov, 1));
  return bytes;
}

// lx: (G+1, N, N) real, the Cholesky vectors followed by h1.  Walkers go through in chunks of wc.
int trot_cisd_energy(const double* gt, const double* bs, const double* ps, const double* omega, const double* lx,
                     const double* D, const void* d_planes, int W, int N, int n, int G, int core, int na, int va,
                     double h0, int wc, double* e, void* ws, size_t ws_bytes, void* stream) {
  if (!gt || !bs || !ps || !omega || !lx || !D || !e) return fail(TROT_ERR_INVALID, "cisd_energy: null pointer");
  if (int rc = check_shape(W, N, n, core, na, va)) return rc;
  if (G < 1 || wc < 1) return fail(TROT_ERR_INVALID, "cisd_energy: bad sizes");
  if (wc > W) wc = W;
  const long ov = (long)na * va;
  if ((long)(G + 1) * N >= (1L << 31) || (long)wc * G >= (1L << 31) || G + 1 > 65535 * 256 || wc > 65535)
    return fail(TROT_ERR_UNSUPPORTED, "cisd_energy: sizes beyond the launch limits");
  const int b3 = d_planes != nullptr;
  if (b3 && (long)wc * G * 4 > (1L << 31)) return fail(TROT_ERR_UNSUPPORTED, "cisd_energy: chunk too large for bf16");
  if (!ws || ws_bytes < trot_cisd_energy_workspace(wc, N, n, G, na, va, b3))
    return fail(TROT_ERR_WORKSPACE, "cisd_energy: workspace");
  Carve cv{(char*)ws, ws_bytes, 0};
  double* gtall = cv.take<double>((size_t)N * wc * n * 2);
  double* R = cv.take<double>((size_t)(G + 1) * N * wc * n * 2);
  double* X = cv.take<double>((size_t)wc * G * ov * 2);
  double* Z = cv.take<double>((size_t)wc * G * ov * 2);
  double* terms = cv.take<double>((size_t)wc * (G + 1) * 4 * 2);
  double* q = cv.take<double>((size_t)wc * G * 2);
  void* xplanes = b3 ? (void*)cv.take<char>(trot_b3_planes_bytes(wc * G, (int)ov, 1)) : nullptr;
  hipStream_t s = (hipStream_t)stream;
  const int TT = n <= 32 ? 32 : 16;
  const size_t terms_lds = (size_t)(n * (n + 1) + 2 * TT * (n + 1) + na * (TT + 1)) * 16;
  HIPCHK(set_lds((const void*)cisd_terms_kernel, terms_lds));
  for (int w0 = 0; w0 < W; w0 += wc) {
    const int c = W - w0 < wc ? W - w0 : wc;
    const long cols = (long)c * n;
    hipLaunchKernelGGL(cisd_permute_kernel, dim3(blocks_of((long)N * cols)), dim3(256), 0, s, gt + 2 * (long)w0 * N * n,
                       N, n, c, gtall);
    HIPCHK(hipGetLastError());
    GemmP p{};  // R = lx gtall: ((G+1) N x N) (N x c n)
    p.A = lx; p.lda = N; p.sA = 0;
    p.B = gtall; p.ldb = cols; p.sB = 0;
    p.C = R; p.ldc = cols; p.sC = 0;
    p.M = (G + 1) * N; p.N = (int)cols; p.K = N; p.ar = 1.0;
    HIPCHK((launch_gemm_f64<64, 64, false, true, true, true, false, EPI_STORE>(p, 1, s)));
    hipLaunchKernelGGL(cisd_terms_kernel, dim3(G + 1, c), dim3(256), terms_lds, s, R, gt + 2 * (long)w0 * N * n,
                       bs + 2 * (long)w0 * ov, ps + 2 * (long)w0 * N * na, N, n, G, core, na, va, c, TT, terms, X);
    HIPCHK(hipGetLastError());
    if (b3) {  // Z = X D on split-BF16 planes (d_planes: D^T, ov rows of K = ov)
      if (int rc = trot_split_bf16(X, 0, 1, 0, ov, c * G, (int)ov, xplanes, stream)) return rc;
      if (int rc = trot_gemm_b3(xplanes, 1, d_planes, c * G, (int)ov, (int)ov, Z, 1, ov, 1.0, 0.0, 0.0, 1, nullptr, 0,
                                stream))
        return rc;
    } else {
      HIPCHK(times_d(X, D, Z, (long)c * G, ov, s));
    }
    hipLaunchKernelGGL(cisd_rowdot_kernel, dim3((unsigned)((long)c * G)), dim3(256), 0, s, Z, X, ov, q);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(cisd_combine_kernel, dim3(c), dim3(256), 0, s, terms, q, omega + 2 * (long)w0, G, h0,
                       e + 2 * (long)w0);
    HIPCHK(hipGetLastError());
  }
  return TROT_OK;
}

}  // extern "C"
