// Local energy with the exchange contractions on 3-pass split-BF16 MFMA
// (trot/meas/rhf.py:41-60, 128-172; trot/meas/uhf.py:205-231).
//
// Per spin s, walker w and Cholesky vector g:
//   f_g = rot_chol_s[g] (n x N) . GT_s[w] (N x n),  c_g = tr f_g,  K_g = sum_pq f_g[p,q] f_g[q,p]
//   E_L = h0 + e1 + 1/2 sum_g [ (c_g^a + c_g^b)^2 - K_g^a - K_g^b ]
// All f_g for all walkers form one GEMM: rows (w, q) of GT^T (complex, K = orbital) against
// rows (g, p) of rot_chol (real), each block padded to 64 rows.  On the wide split-BF16 tile
// (128 x 256 = 2 walkers x 4 Cholesky vectors) every wave's 64 x 64 sub-tile is exactly one
// f_g^T of one walker: the epilogue parks it in LDS, takes its trace and sum_pq f_pq f_qp in
// FP64, and writes (c_g, K_g) for that (w, g).  f never reaches HBM.  The sum over g and the
// spin combination run in FP64 in a fixed order.
#include "capi_util.hpp"
#include "common.hpp"
#include "gemm_b3.hpp"
#include "kernels.hpp"

namespace {

inline int pad_up(int x, int m) { return (x + m - 1) / m * m; }

// GT (W, N, n) complex128 -> planes [q][w * 64 + j][i] (q: re_hi, re_lo, im_hi, im_lo), Kp = Np, rows
// padded to rows_pad (multiple of 128), zero beyond (W, N, n).  Grid (Np / 64, rows_pad / 64).
__global__ __launch_bounds__(256) void split_gt_b3_kernel(const double* __restrict__ gt, bf16raw* __restrict__ out,
                                                          int W, int N, int n, int Np, long rows_pad) {
  __shared__ double2 tile[64][65];
  const int i0 = blockIdx.x * 64;
  const int w = blockIdx.y;
  const int t = threadIdx.x;
  const double2* gw = reinterpret_cast<const double2*>(gt) + (long)w * N * n;
  for (int e = t; e < 64 * 64; e += 256) {
    const int ii = e >> 6, jj = e & 63;  // read along j
    const int gi = i0 + ii;
    tile[ii][jj] = (w < W && gi < N && jj < n) ? gw[(long)gi * n + jj] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const long plane = rows_pad * Np;
  for (int e = t; e < 64 * 64; e += 256) {
    const int jj = e >> 6, ii = e & 63;  // write along i
    const double2 v = tile[ii][jj];
    const long o = ((long)w * 64 + jj) * Np + i0 + ii;
    bf16raw h, l;
    split_bf16_raw((float)v.x, h, l);
    out[o] = h;
    out[plane + o] = l;
    split_bf16_raw((float)v.y, h, l);
    out[2 * plane + o] = h;
    out[3 * plane + o] = l;
  }
}

// rot_chol (G, n, N) FP64 -> planes [q][g * 64 + p][i] (hi, lo), rows padded to rows_pad, Kp = Np.
__global__ void split_rot_chol_b3_kernel(const double* __restrict__ rc, bf16raw* __restrict__ out, int G, int n,
                                         int N, int Np, long rows_pad) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = rows_pad * Np;
  if (e >= total) return;
  const long row = e / Np;
  const int i = (int)(e - row * Np);
  const long g = row >> 6;
  const int p = (int)(row & 63);
  const double v = (g < G && p < n && i < N) ? rc[(g * n + p) * N + i] : 0.0;
  bf16raw h, l;
  split_bf16_raw((float)v, h, l);
  out[e] = h;
  out[total + e] = l;
}

// ck[(w G + g) * 4 + {0,1,2,3}] = (Re c_g, Im c_g, Re K_g, Im K_g) for one spin.
__global__ __launch_bounds__(512, 1) void exchange_b3_kernel(B3P p, int n_mt, int n_nt, int b_rows, int W, int G,
                                                             double* __restrict__ ck) {
  __shared__ __attribute__((aligned(16))) char smem[B3W_SMEM];
  int mt, nt;
  if (!b3_tile_of(blockIdx.x, n_mt, n_nt, 8, mt, nt)) return;
  f4_t acc[2][4][4];
  b3w_mainloop(p, mt * B3W_BM, nt * B3W_BN, b_rows, smem, acc, blockIdx.y);  // ends on a barrier: smem is free

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int w = mt * 2 + (wave >> 2), g = nt * 4 + (wave & 3);
  // X = f_g (64 x 64 complex, FP32, re and im planes) of this wave, element (r, c) at r * 64 + (c ^ r); four
  // waves at a time
  // (4 x 32 KB = the main loop's LDS).  Each lane holds 64 elements f_ij of its wave's f_g in registers
  // (row 16 a + 4 (lane / 16) + r, column 16 b + lane % 16) and reads only their partners f_ji from X:
  // K_g = sum_ij f_ij f_ji visits every ordered pair once; c_g = tr f_g from the diagonal ones.  Each product
  // is formed in FP32 (one rounding, 2^-24, far below the split-BF16 error of f) and summed in FP64.
  float* Xr = reinterpret_cast<float*>(smem) + (wave & 3) * 8192;  // re plane, im plane 4096 floats further
  float* Xi = Xr + 4096;
  // row 16 i + 4 q + r, column 16 j + c (q = lane / 16, c = lane % 16): (row, col) sits at
  // 1024 i + 16 (i ^ j) + wr[r] and its partner (col, row) at 1024 j + 16 (i ^ j) + pr[r]
  const int q = lane >> 4, c = lane & 15;
  int wr[4], pr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    wr[r] = 64 * (4 * q + r) + (c ^ (4 * q + r));
    pr[r] = 64 * c + (c ^ (4 * q + r));
  }
  for (int round = 0; round < 2; ++round) {
    const bool mine = (wave >> 2) == round;
    if (mine) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int o = 1024 * i + 16 * (i ^ j) + wr[r];
            Xr[o] = acc[0][i][j][r];
            Xi[o] = acc[1][i][j][r];
          }
    }
    __syncthreads();
    if (mine) {
      double cr = 0.0, ci = 0.0, kr = 0.0, ki = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int o = 1024 * j + 16 * (i ^ j) + pr[r];
            const float xr = acc[0][i][j][r], xi = acc[1][i][j][r];
            const float tr = Xr[o], ti = Xi[o];  // f_ji
            kr += (double)(xr * tr - xi * ti);
            ki += (double)(xr * ti + xi * tr);
            if (i == j && c == 4 * q + r) {  // diagonal
              cr += xr;
              ci += xi;
            }
          }
        __builtin_amdgcn_sched_barrier(0);  // one row tile's reads in flight at a time
      }
#pragma unroll
      for (int m = 32; m > 0; m >>= 1) {
        cr += __shfl_xor(cr, m);
        ci += __shfl_xor(ci, m);
        kr += __shfl_xor(kr, m);
        ki += __shfl_xor(ki, m);
      }
      if (lane == 0 && w < W && g < G) {
        double2* o = reinterpret_cast<double2*>(ck + ((long)w * G + g) * 4);
        o[0] = make_double2(cr, ci);
        o[1] = make_double2(kr, ki);
      }
    }
    __syncthreads();
  }
}

// partial[w] = 1/2 sum_g [ (c_a + c_b)^2 - K_a - K_b ], second spin = first when ckb is null.
__global__ __launch_bounds__(256) void exchange_combine_kernel(const double* __restrict__ cka,
                                                               const double* __restrict__ ckb, int G,
                                                               double* __restrict__ partial) {
  __shared__ double red[256];
  const long w = blockIdx.x;
  const int t = threadIdx.x;
  double sr = 0.0, si = 0.0;
  for (int g = t; g < G; g += 256) {
    const double* a = cka + (w * G + g) * 4;
    const double* b = ckb ? ckb + (w * G + g) * 4 : a;
    const cplx cs = {a[0] + b[0], a[1] + b[1]};
    const cplx c2 = cmul(cs, cs);
    sr += c2.re - a[2] - b[2];
    si += c2.im - a[3] - b[3];
  }
  sr = block_sum<256>(sr, red);
  si = block_sum<256>(si, red);
  if (t == 0) cstore(partial, w, {0.5 * sr, 0.5 * si});
}

struct ExLayout {
  int Np;
  long a_rows, b_rows;
};

ExLayout ex_layout(int W, int N, int G) {
  ExLayout L;
  L.Np = pad_up(N, 128);
  L.a_rows = (long)pad_up(W, 2) * 64;
  L.b_rows = (long)pad_up(G, 4) * 64;
  return L;
}

int exchange_spin(const bf16raw* rcp, const double* gt, int W, int N, int n, int G, bf16raw* ap, double* ck,
                  hipStream_t s) {
  const ExLayout L = ex_layout(W, N, G);
  hipLaunchKernelGGL(split_gt_b3_kernel, dim3(L.Np / 64, (unsigned)(L.a_rows / 64)), dim3(256), 0, s, gt, ap, W, N, n,
                     L.Np, L.a_rows);
  HIPCHK(hipGetLastError());
  B3P p{};
  p.A = ap;
  p.B = rcp;
  p.a_plane = L.a_rows * L.Np;
  p.b_plane = L.b_rows * L.Np;
  p.Kp = L.Np;
  p.k_slice = L.Np;
  p.M = (int)L.a_rows;
  p.N = (int)L.b_rows;
  const int n_mt = (int)(L.a_rows / B3W_BM), n_nt = (int)(L.b_rows / B3W_BN);
  const long T = (long)n_mt * n_nt;
  if (T > 0x7fffffffL / 2) return fail(TROT_ERR_INVALID, "local_energy_b3: too many tiles");
  hipLaunchKernelGGL(exchange_b3_kernel, dim3((unsigned)((T + 7) / 8 * 8)), dim3(512), 0, s, p, n_mt, n_nt,
                     (int)L.b_rows, W, G, ck);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

}  // namespace

extern "C" {

size_t trot_rot_chol_b3_bytes(int G, int n, int N) {
  const ExLayout L = ex_layout(1, N, G < 1 ? 1 : G);
  (void)n;
  return (size_t)2 * L.b_rows * L.Np * sizeof(bf16raw);
}

int trot_rot_chol_b3(const double* rot_chol, int G, int n, int N, void* planes, void* stream) {
  if (!rot_chol || !planes) return fail(TROT_ERR_INVALID, "rot_chol_b3: null pointer");
  if (G < 1 || n < 1 || N < 1) return fail(TROT_ERR_INVALID, "rot_chol_b3: bad sizes");
  if (n > 64) return fail(TROT_ERR_UNSUPPORTED, "rot_chol_b3: nocc > 64");
  const ExLayout L = ex_layout(1, N, G);
  const long total = L.b_rows * L.Np;
  hipLaunchKernelGGL(split_rot_chol_b3_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, rot_chol, (bf16raw*)planes, G, n, N, L.Np, L.b_rows);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

size_t trot_local_energy_b3_workspace(int W, int N, int G, int two_spins) {
  const ExLayout L = ex_layout(W, N, G < 1 ? 1 : G);
  const size_t ck = align256((size_t)W * (G < 1 ? 1 : G) * 32);
  return align256((size_t)4 * L.a_rows * L.Np * sizeof(bf16raw)) + (two_spins ? 2 : 1) * ck + align256((size_t)W * 16);
}

int trot_local_energy_b3(const void* rc_planes_a, const void* rc_planes_b, const double* green_t_a,
                         const double* green_t_b, const double* rot_h1t_a, const double* rot_h1t_b, int W, int N,
                         int na, int nb, int G, double h0, double* energy, void* ws, size_t ws_bytes, void* stream) {
  if (!green_t_a || !rot_h1t_a || !energy || (G > 0 && !rc_planes_a))
    return fail(TROT_ERR_INVALID, "local_energy_b3: null pointer");
  if (green_t_b && (!rot_h1t_b || (G > 0 && !rc_planes_b)))
    return fail(TROT_ERR_INVALID, "local_energy_b3: second spin incomplete");
  if (W < 1 || N < 1 || na < 1 || G < 0 || (green_t_b && nb < 1))
    return fail(TROT_ERR_INVALID, "local_energy_b3: bad sizes");
  if (na > 64 || (green_t_b && nb > 64)) return fail(TROT_ERR_UNSUPPORTED, "local_energy_b3: nocc > 64");
  if (W > 65534) return fail(TROT_ERR_UNSUPPORTED, "local_energy_b3: more than 65534 walkers per call");
  const bool two = green_t_b != nullptr;
  if (ws_bytes < trot_local_energy_b3_workspace(W, N, G, two))
    return fail(TROT_ERR_WORKSPACE, "local_energy_b3: workspace");
  hipStream_t s = (hipStream_t)stream;
  const ExLayout L = ex_layout(W, N, G < 1 ? 1 : G);
  Carve cv{(char*)ws, ws_bytes, 0};
  bf16raw* ap = cv.take<bf16raw>((size_t)4 * L.a_rows * L.Np);
  double* cka = cv.take<double>((size_t)W * (G < 1 ? 1 : G) * 4);
  double* ckb = two ? cv.take<double>((size_t)W * (G < 1 ? 1 : G) * 4) : nullptr;
  double* partial = cv.take<double>((size_t)W * 2);
  if (!ap || !cka || (two && !ckb) || !partial) return fail(TROT_ERR_WORKSPACE, "local_energy_b3: workspace");
  int nblk = 0;
  if (G > 0) {
    if (int rc = exchange_spin((const bf16raw*)rc_planes_a, green_t_a, W, N, na, G, ap, cka, s)) return rc;
    if (two)
      if (int rc = exchange_spin((const bf16raw*)rc_planes_b, green_t_b, W, N, nb, G, ap, ckb, s)) return rc;
    hipLaunchKernelGGL(exchange_combine_kernel, dim3(W), dim3(256), 0, s, cka, ckb, G, partial);
    HIPCHK(hipGetLastError());
    nblk = 1;
  }
  HIPCHK(launch_energy_finalize(green_t_a, green_t_b, rot_h1t_a, rot_h1t_b, na, nb, N, h0, partial, nblk, energy, W,
                                s));
  return TROT_OK;
}

}  // extern "C"
