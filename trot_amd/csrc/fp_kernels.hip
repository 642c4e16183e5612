// Free-projection step and block (trot/prop/afqmc_fp.py:18-72, trot/prop/blocks.py:415-490).
//
//   fp_prepare_kernel      shift term, walker constant and its per-orbital root, unshifted xbar
//                          (afqmc_fp.py:91-96 with multiply_constants, trot/walkers.py:208-226)
//   scale_rows_kernel      w[b] *= s[b]
//   fp_weights_kernel      QR norms from |det R| per spin, weights * norm (afqmc_fp.py:107-108,
//                          trot/walkers.py:182-205)
//   fp_unnorm_kernel       weights after the comb divided by the norm of the walker they copy
//                          (afqmc_fp.py:113-119)
//   block_energy_fp_kernel clamp around ene0 and the overlap-weighted block sums (blocks.py:466-488)
#include "capi_util.hpp"
#include "common.hpp"

namespace {

// One workgroup per walker.  z = -sqrt(dt) sum_g x_g mf_g + dt h0p; c = exp(z); s = c^(1 / n_total)
// by the principal logarithm of c (as a complex power does); xbar = x + 0i.
__global__ __launch_bounds__(256) void fp_prepare_kernel(const double* __restrict__ fields,
                                                         const double* __restrict__ mf, int G, double sqrt_dt,
                                                         double dt, double h0p_re, double h0p_im, double inv_n,
                                                         double* __restrict__ xbar, double* __restrict__ cst,
                                                         double* __restrict__ scale) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  const long w = blockIdx.x;
  double s_re = 0.0, s_im = 0.0;
  for (int g = t; g < G; g += 256) {
    const double x = fields[w * G + g];
    const cplx m = cload(mf, g);
    s_re += x * m.re;
    s_im += x * m.im;
    if (xbar) cstore(xbar, w * G + g, cplx{x, 0.0});
  }
  s_re = block_sum<256>(s_re, red);
  s_im = block_sum<256>(s_im, red);
  if (t == 0) {
    const double zr = -sqrt_dt * s_re + dt * h0p_re, zi = -sqrt_dt * s_im + dt * h0p_im;
    const double ez = exp(zr);
    const cplx c = {ez * cos(zi), ez * sin(zi)};
    if (cst) cstore(cst, w, c);
    const double lr = log(hypot(c.re, c.im)) * inv_n, th = atan2(c.im, c.re) * inv_n;
    const double m = exp(lr);
    cstore(scale, w, cplx{m * cos(th), m * sin(th)});
  }
}

__global__ __launch_bounds__(256) void scale_rows_kernel(double* __restrict__ a, const double* __restrict__ s,
                                                         long row) {
  const long b = blockIdx.x;
  const cplx f = cload(s, b);
  double2* r = reinterpret_cast<double2*>(a) + b * row;
  for (long e = threadIdx.x; e < row; e += 256) {
    const double2 v = r[e];
    r[e] = make_double2(v.x * f.re - v.y * f.im, v.x * f.im + v.y * f.re);
  }
}

// norm = det_a^2 (restricted walkers) or det_a det_b; weights_new = weights * norm
__global__ void fp_weights_kernel(const double* __restrict__ weights, const double* __restrict__ det_a,
                                  const double* __restrict__ det_b, int W, double* __restrict__ norms,
                                  double* __restrict__ weights_new) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const double a = det_a[w];
  const double nrm = det_b ? a * det_b[w] : a * a;
  norms[w] = nrm;
  weights_new[w] = weights[w] * nrm;
}

__global__ void fp_unnorm_kernel(double* __restrict__ weights, const double* __restrict__ norms,
                                 const int* __restrict__ idx, int W) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  weights[w] /= norms[idx[w]];
}

// e (complex E_L), weights (real), overlaps (complex).  e -> ene0 where |e - ene0| > thresh;
// out = [sum w ov e / sum w ov, sum w ov, sum ov, sum |ov|] as complex pairs.
__global__ __launch_bounds__(1024) void block_energy_fp_kernel(const double* __restrict__ energy,
                                                               const double* __restrict__ weights,
                                                               const double* __restrict__ ovlp, int W, double thresh,
                                                               double ene0, double* __restrict__ out) {
  __shared__ double red[1024];
  const int t = threadIdx.x;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};  // wov re/im, wove re/im, ov re/im, |ov|
  for (int w = t; w < W; w += 1024) {
    cplx e = cload(energy, w);
    if (hypot(e.re - ene0, e.im) > thresh) e = {ene0, 0.0};
    const cplx ov = cload(ovlp, w);
    const cplx wov = cscale(ov, weights[w]);
    const cplx wove = cmul(wov, e);
    acc[0] += wov.re;
    acc[1] += wov.im;
    acc[2] += wove.re;
    acc[3] += wove.im;
    acc[4] += ov.re;
    acc[5] += ov.im;
    acc[6] += hypot(ov.re, ov.im);
  }
#pragma unroll
  for (int q = 0; q < 7; ++q) acc[q] = block_sum<1024>(acc[q], red);
  if (t == 0) {
    const cplx wov = {acc[0], acc[1]}, wove = {acc[2], acc[3]};
    const cplx eb = cmul(wove, cinv(wov));
    out[0] = eb.re;
    out[1] = eb.im;
    out[2] = wov.re;
    out[3] = wov.im;
    out[4] = acc[4];
    out[5] = acc[5];
    out[6] = acc[6];
    out[7] = 0.0;
  }
}

}  // namespace

extern "C" {

int trot_fp_prepare(const double* fields, const double* mf, int W, int G, double sqrt_dt, double dt, double h0p_re,
                    double h0p_im, int n_total, double* xbar, double* cst, double* scale, void* stream) {
  if (!fields || !mf || !scale) return fail(TROT_ERR_INVALID, "fp_prepare: null pointer");
  if (W < 1 || G < 1 || n_total < 1) return fail(TROT_ERR_INVALID, "fp_prepare: bad sizes");
  hipLaunchKernelGGL(fp_prepare_kernel, dim3(W), dim3(256), 0, (hipStream_t)stream, fields, mf, G, sqrt_dt, dt,
                     h0p_re, h0p_im, 1.0 / n_total, xbar, cst, scale);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_scale_rows(double* a, const double* scale, int W, long row_complex, void* stream) {
  if (!a || !scale || W < 0 || row_complex < 0) return fail(TROT_ERR_INVALID, "scale_rows: bad arguments");
  if (W == 0 || row_complex == 0) return TROT_OK;
  hipLaunchKernelGGL(scale_rows_kernel, dim3(W), dim3(256), 0, (hipStream_t)stream, a, scale, row_complex);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_fp_weights(const double* weights, const double* det_a, const double* det_b, int W, double* norms,
                    double* weights_new, void* stream) {
  if (!weights || !det_a || !norms || !weights_new || W < 1) return fail(TROT_ERR_INVALID, "fp_weights: bad arguments");
  hipLaunchKernelGGL(fp_weights_kernel, dim3((W + 255) / 256), dim3(256), 0, (hipStream_t)stream, weights, det_a,
                     det_b, W, norms, weights_new);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_fp_unnorm(double* weights, const double* norms, const int* idx, int W, void* stream) {
  if (!weights || !norms || !idx || W < 1) return fail(TROT_ERR_INVALID, "fp_unnorm: bad arguments");
  hipLaunchKernelGGL(fp_unnorm_kernel, dim3((W + 255) / 256), dim3(256), 0, (hipStream_t)stream, weights, norms, idx,
                     W);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_block_energy_fp(const double* energy, const double* weights, const double* overlaps, int W, double dt,
                         double ene0, double* out, void* stream) {
  if (!energy || !weights || !overlaps || !out || W < 1 || !(dt > 0.0))
    return fail(TROT_ERR_INVALID, "block_energy_fp: bad arguments");
  hipLaunchKernelGGL(block_energy_fp_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, energy, weights, overlaps, W,
                     sqrt(2.0 / dt), ene0, out);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

}  // extern "C"
