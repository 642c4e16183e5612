// Block observables: per-walker one-body density matrices and density correlations, and their
// weighted mean over the walkers of a block.
//
//   rdm1_kernel          dm[w, s] = GT_s[w] C_s^T: the (N x n) transposed half Green's function times the
//                        (n x N) trial transpose, per walker and spin (trot/meas/rhf.py:191-216,
//                        trot/meas/uhf.py:97-110)
//   density_corr_kernel  (uu, ud, dd)[i, j] from the two spins' dm (trot/meas/uhf.py:129-144)
//   wmean_partial_kernel / wmean_final_kernel
//                        out = sum_w weight_w x_w / sum_w weight_w, 0 when the weight sum is 0
//                        (trot/prop/blocks.py:259-268), summed in a fixed order (walker slices, then
//                        the slices in order) so that every call gives the same bits
#include "capi_util.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int RDM_TP = 16;  // output rows per workgroup (each thread keeps RDM_TP complex sums)

// grid (ceil(N / blockDim.x), ceil(N / RDM_TP), W).  Thread t owns column q; the workgroup's RDM_TP rows
// of GT sit in LDS and are read as broadcasts.
__global__ __launch_bounds__(256) void rdm1_kernel(const double* __restrict__ gt, const double* __restrict__ c,
                                                   int N, int n, int S, int spin, double* __restrict__ out) {
  __shared__ double g_lds[RDM_TP * TROT_MAX_NOCC * 2];
  const long w = blockIdx.z;
  const int p0 = blockIdx.y * RDM_TP;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int rows = min(RDM_TP, N - p0);
  const double* gw = gt + (w * N + p0) * (long)n * 2;
  for (int e = threadIdx.x; e < rows * n * 2; e += blockDim.x) g_lds[e] = gw[e];
  __syncthreads();
  if (q >= N) return;
  double acc_re[RDM_TP], acc_im[RDM_TP];
#pragma unroll
  for (int p = 0; p < RDM_TP; ++p) acc_re[p] = acc_im[p] = 0.0;
  const double* cq = c + (long)q * n;
  for (int i = 0; i < n; ++i) {
    const double ci = cq[i];
#pragma unroll
    for (int p = 0; p < RDM_TP; ++p) {
      if (p < rows) {
        acc_re[p] += g_lds[(p * n + i) * 2] * ci;
        acc_im[p] += g_lds[(p * n + i) * 2 + 1] * ci;
      }
    }
  }
  double* o = out + ((w * S + spin) * N + p0) * (long)N * 2;
#pragma unroll
  for (int p = 0; p < RDM_TP; ++p)
    if (p < rows) cstore(o, (long)p * N + q, cplx{acc_re[p], acc_im[p]});
}

// dm (W, 2, N, N) -> out (W, 3, N, N):
//   uu = na_i na_j - ga_ij ga_ji + delta_ij na_i, ud = na_i nb_j, dd as uu with gb; n = diag(g).
__global__ __launch_bounds__(256) void density_corr_kernel(const double* __restrict__ dm, int N, long total,
                                                           double* __restrict__ out) {
  const long NN = (long)N * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long w = e / NN;
    const int ij = (int)(e - w * NN);
    const int i = ij / N, j = ij - i * N;
    const double* ga = dm + w * 2 * NN * 2;
    const double* gb = ga + NN * 2;
    const cplx nai = cload(ga, (long)i * N + i), naj = cload(ga, (long)j * N + j);
    const cplx nbi = cload(gb, (long)i * N + i), nbj = cload(gb, (long)j * N + j);
    cplx uu = cmul(nai, naj), dd = cmul(nbi, nbj);
    const cplx xa = cmul(cload(ga, ij), cload(ga, (long)j * N + i));
    const cplx xb = cmul(cload(gb, ij), cload(gb, (long)j * N + i));
    uu.re -= xa.re;
    uu.im -= xa.im;
    dd.re -= xb.re;
    dd.im -= xb.im;
    if (i == j) {
      uu.re += nai.re;
      uu.im += nai.im;
      dd.re += nbi.re;
      dd.im += nbi.im;
    }
    double* o = out + w * 3 * NN * 2;
    cstore(o, ij, uu);
    cstore(o, NN + ij, cmul(nai, nbj));
    cstore(o, 2 * NN + ij, dd);
  }
}

// partial[y, col] = sum over rows r in slice y (in order) of weights[r] x[r, col].
__global__ __launch_bounds__(256) void wmean_partial_kernel(const double* __restrict__ x,
                                                            const double* __restrict__ weights, int W, long L,
                                                            int rows_per_slice, double* __restrict__ partial) {
  const long col = blockIdx.x * 256L + threadIdx.x;
  if (col >= L) return;
  const int r0 = blockIdx.y * rows_per_slice;
  const int r1 = min(W, r0 + rows_per_slice);
  double s = 0.0;
  for (int r = r0; r < r1; ++r) s += weights[r] * x[r * L + col];
  partial[blockIdx.y * L + col] = s;
}

__global__ __launch_bounds__(256) void wmean_final_kernel(const double* __restrict__ partial, int slices, long L,
                                                          const double* __restrict__ wsum, double* __restrict__ out) {
  const long col = blockIdx.x * 256L + threadIdx.x;
  if (col >= L) return;
  double s = 0.0;
  for (int y = 0; y < slices; ++y) s += partial[y * L + col];
  const double ws = *wsum;
  out[col] = ws == 0.0 ? 0.0 : s / ws;
}

// Walker slices of the weighted mean: enough workgroups to fill the chip (about 2048), each slice at
// least one walker.
void wmean_shape(int W, long L, int* slices, int* rows) {
  const long bx = (L + 255) / 256;
  long s = 2048 / (bx > 0 ? bx : 1);
  if (s < 1) s = 1;
  if (s > W) s = W;
  const int r = (int)((W + s - 1) / s);
  *rows = r;
  *slices = (W + r - 1) / r;
}

}  // namespace

extern "C" {

int trot_rdm1_from_green(const double* green_t, const double* trial, int W, int N, int n, int n_spin_out, int spin,
                         double* out, void* stream) {
  if (!green_t || !trial || !out) return fail(TROT_ERR_INVALID, "rdm1_from_green: null pointer");
  if (W < 1 || N < 1 || n < 1 || n > N || n_spin_out < 1 || spin < 0 || spin >= n_spin_out)
    return fail(TROT_ERR_INVALID, "rdm1_from_green: bad sizes");
  if (n > TROT_MAX_NOCC) return fail(TROT_ERR_UNSUPPORTED, "rdm1_from_green: nocc > 64");
  if (W > 65535) return fail(TROT_ERR_UNSUPPORTED, "rdm1_from_green: more than 65535 walkers in one call");
  const int bx = N >= 256 ? 256 : ((N + 63) / 64) * 64;
  dim3 grid((N + bx - 1) / bx, (N + RDM_TP - 1) / RDM_TP, W);
  hipLaunchKernelGGL(rdm1_kernel, grid, dim3(bx), 0, (hipStream_t)stream, green_t, trial, N, n, n_spin_out, spin,
                     out);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_density_corr(const double* dm, int W, int N, double* out, void* stream) {
  if (!dm || !out) return fail(TROT_ERR_INVALID, "density_corr: null pointer");
  if (W < 1 || N < 1) return fail(TROT_ERR_INVALID, "density_corr: bad sizes");
  const long total = (long)W * N * N;
  long blocks = (total + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(density_corr_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, dm, N, total,
                     out);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

size_t trot_weighted_mean_workspace(int W, long L) {
  if (W < 1 || L < 1) return 0;
  int slices, rows;
  wmean_shape(W, L, &slices, &rows);
  return align256((size_t)slices * L * sizeof(double));
}

int trot_weighted_mean(const double* x, const double* weights, const double* wsum, int W, long L, double* out,
                       void* ws, size_t ws_bytes, void* stream) {
  if (!x || !weights || !wsum || !out || !ws) return fail(TROT_ERR_INVALID, "weighted_mean: null pointer");
  if (W < 1 || L < 1) return fail(TROT_ERR_INVALID, "weighted_mean: bad sizes");
  if (ws_bytes < trot_weighted_mean_workspace(W, L)) return fail(TROT_ERR_INVALID, "weighted_mean: workspace too small");
  int slices, rows;
  wmean_shape(W, L, &slices, &rows);
  const long bx = (L + 255) / 256;
  if (bx > 0x7fffffffL) return fail(TROT_ERR_UNSUPPORTED, "weighted_mean: too many columns");
  hipLaunchKernelGGL(wmean_partial_kernel, dim3((unsigned)bx, slices), dim3(256), 0, (hipStream_t)stream, x, weights,
                     W, L, rows, (double*)ws);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(wmean_final_kernel, dim3((unsigned)bx), dim3(256), 0, (hipStream_t)stream, (const double*)ws,
                     slices, L, wsum, out);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

}  // extern "C"
