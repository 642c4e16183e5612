// Local energy (trot/meas/rhf.py:41-60, 128-172; trot/meas/uhf.py:205-231).
//
// E_L = h0 + e1 + 1/2 sum_g [ (c_g^a + c_g^b)^2 - K_g^a - K_g^b ],
//   f_g^s = rot_chol_s[g] (n_s x N) . GT_s (N x n_s)       (GT = G_half^T)
//   c_g^s = tr f_g^s,  K_g^s = sum_pq f_g^s[p,q] f_g^s[q,p]
// With one spin (RHF trial, restricted walkers) both spins are equal:
// c = 2 c^a, K = 2 K^a, which is 2 sum c^2 - sum K of the reference.
//
// exchange_kernel: grid (ceil(G / GB), W).  Each workgroup takes one walker
// and GB Cholesky vectors; every f_g tile (<= 64 x 64; exchange_big_kernel: 64 x 64 tile pairs for
// nocc up to 256) is produced on MFMA
// (FP64, FP32 or split-BF16), parked in LDS for the transpose, reduced in
// FP64, and dropped: f is never written to HBM (the reference's "high" mode
// materialises all of it, "low" scans g; both give the same sums).
#include "common.hpp"
#include "gemm.hpp"
#include "gemm_lp.hpp"
#include "kernels.hpp"

typedef TileF64<64, 64, false, true, true, true, false> ExTile64;
typedef TileLP<64, 64, MODE_F32, false, true, float, double, true, true, false> ExTile32;
typedef TileLP<64, 64, MODE_BF16X3, false, true, float, double, true, true, false> ExTileB3;

template <typename Tile>
struct ExOps;
template <>
struct ExOps<ExTile64> {
  typedef double TA;
};
template <>
struct ExOps<ExTile32> {
  typedef float TA;
};
template <>
struct ExOps<ExTileB3> {
  typedef float TA;
};

template <typename Tile>
__device__ __forceinline__ void exchange_one(const typename ExOps<Tile>::TA* __restrict__ rc,
                                             const double* __restrict__ gt, int n, int N, typename Tile::Smem& sm,
                                             cplx* F, double* red, cplx& c, cplx& k) {
  typename Tile::Acc acc;
  Tile::run(rc, N, gt, n, n, n, N, 0, 0, sm, acc);
  __syncthreads();  // sm is aliased by F
#pragma unroll
  for (int i = 0; i < Tile::TM; ++i)
#pragma unroll
    for (int j = 0; j < Tile::TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int p = Tile::row_of(i, r), q = Tile::col_of(j);
        F[p * 64 + q] = {(double)acc[0][i][j][r], (double)acc[1][i][j][r]};
      }
  __syncthreads();
  const int t = threadIdx.x;
  double cr = 0, ci = 0, kr = 0, ki = 0;
  if (t < n) {
    cr = F[t * 64 + t].re;
    ci = F[t * 64 + t].im;
  }
  for (int e = t; e < n * n; e += 256) {
    const int p = e / n, q = e - (e / n) * n;
    const cplx v = cmul(F[p * 64 + q], F[q * 64 + p]);
    kr += v.re;
    ki += v.im;
  }
  cr = block_sum<256>(cr, red);
  ci = block_sum<256>(ci, red);
  kr = block_sum<256>(kr, red);
  ki = block_sum<256>(ki, red);
  c = {cr, ci};
  k = {kr, ki};
}

template <typename Tile>
__global__ __launch_bounds__(256) void exchange_kernel(const typename ExOps<Tile>::TA* __restrict__ rc_a,
                                                       const typename ExOps<Tile>::TA* __restrict__ rc_b,
                                                       const double* __restrict__ gt_a, const double* __restrict__ gt_b,
                                                       int na, int nb, int N, int G, int GB,
                                                       double* __restrict__ partial) {
  __shared__ union {
    typename Tile::Smem sm;
    cplx F[64 * 64];
  } u;
  __shared__ double red[256];
  const long w = blockIdx.y;
  const int gblk = blockIdx.x;
  const int g0 = gblk * GB;
  const int g1 = min(G, g0 + GB);
  const double* ga = gt_a + 2 * w * (long)N * na;
  const double* gb = gt_b ? gt_b + 2 * w * (long)N * nb : nullptr;
  double sr = 0, si = 0;
  for (int g = g0; g < g1; ++g) {
    cplx ca, ka, cb, kb;
    exchange_one<Tile>(rc_a + (long)g * na * N, ga, na, N, u.sm, u.F, red, ca, ka);
    if (gb) {
      exchange_one<Tile>(rc_b + (long)g * nb * N, gb, nb, N, u.sm, u.F, red, cb, kb);
    } else {
      cb = ca;
      kb = ka;
    }
    const cplx cs = cadd(ca, cb);
    const cplx c2 = cmul(cs, cs);
    sr += c2.re - ka.re - kb.re;
    si += c2.im - ka.im - kb.im;
  }
  if (threadIdx.x == 0) cstore(partial, w * gridDim.x + gblk, {0.5 * sr, 0.5 * si});
}

// nocc > 64: f_g in 64 x 64 tiles.  c = sum_P tr f_PP; K = sum_P sum f_PP o f_PP^T + 2 sum_{P<Q} sum f_PQ o f_QP^T
// (the (P, Q) and (Q, P) terms of sum_pq f_pq f_qp are equal); tile (P, Q) in F1, tile (Q, P) in F2 (aliasing
// the GEMM staging).  Every sum runs in a fixed order.
template <typename Tile>
__device__ __forceinline__ void exchange_one_big(const typename ExOps<Tile>::TA* __restrict__ rc,
                                                 const double* __restrict__ gt, int n, int N, typename Tile::Smem& sm,
                                                 cplx* F2, cplx* F1, double* red, cplx& c, cplx& k) {
  const int nt = (n + 63) / 64;
  const int t = threadIdx.x;
  double cr = 0, ci = 0, kr = 0, ki = 0;
  for (int P = 0; P < nt; ++P)
    for (int Q = P; Q < nt; ++Q) {
      typename Tile::Acc acc;
      for (int pass = 0; pass < (P == Q ? 1 : 2); ++pass) {
        const int r0 = 64 * (pass ? Q : P), c0 = 64 * (pass ? P : Q);
        Tile::run(rc, N, gt, n, n, n, N, r0, c0, sm, acc);
        __syncthreads();  // sm is aliased by F2
        cplx* F = pass ? F2 : F1;
#pragma unroll
        for (int i = 0; i < Tile::TM; ++i)
#pragma unroll
          for (int j = 0; j < Tile::TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int pp = Tile::row_of(i, r), qq = Tile::col_of(j);
              F[pp * 64 + qq] = {(double)acc[0][i][j][r], (double)acc[1][i][j][r]};
            }
        __syncthreads();
      }
      const int rp = min(64, n - 64 * P), rq = min(64, n - 64 * Q);
      if (P == Q) {
        if (t < rp) {
          cr += F1[t * 64 + t].re;
          ci += F1[t * 64 + t].im;
        }
        for (int e = t; e < rp * rp; e += 256) {
          const int pp = e / rp, qq = e - (e / rp) * rp;
          const cplx v = cmul(F1[pp * 64 + qq], F1[qq * 64 + pp]);
          kr += v.re;
          ki += v.im;
        }
      } else {
        for (int e = t; e < rp * rq; e += 256) {
          const int pp = e / rq, qq = e - (e / rq) * rq;
          const cplx v = cmul(F1[pp * 64 + qq], F2[qq * 64 + pp]);  // f[64P + pp, 64Q + qq] f[64Q + qq, 64P + pp]
          kr += 2.0 * v.re;
          ki += 2.0 * v.im;
        }
      }
      __syncthreads();  // F1 / F2 / sm are rewritten by the next tile pair
    }
  cr = block_sum<256>(cr, red);
  ci = block_sum<256>(ci, red);
  kr = block_sum<256>(kr, red);
  ki = block_sum<256>(ki, red);
  c = {cr, ci};
  k = {kr, ki};
}

template <typename Tile>
__global__ __launch_bounds__(256) void exchange_big_kernel(const typename ExOps<Tile>::TA* __restrict__ rc_a,
                                                           const typename ExOps<Tile>::TA* __restrict__ rc_b,
                                                           const double* __restrict__ gt_a,
                                                           const double* __restrict__ gt_b, int na, int nb, int N,
                                                           int G, int GB, double* __restrict__ partial) {
  __shared__ union {
    typename Tile::Smem sm;
    cplx F[64 * 64];
  } u;
  __shared__ cplx F1[64 * 64];
  __shared__ double red[256];
  const long w = blockIdx.y;
  const int gblk = blockIdx.x;
  const int g0 = gblk * GB;
  const int g1 = min(G, g0 + GB);
  const double* ga = gt_a + 2 * w * (long)N * na;
  const double* gb = gt_b ? gt_b + 2 * w * (long)N * nb : nullptr;
  double sr = 0, si = 0;
  for (int g = g0; g < g1; ++g) {
    cplx ca, ka, cb, kb;
    if (na > 64) exchange_one_big<Tile>(rc_a + (long)g * na * N, ga, na, N, u.sm, u.F, F1, red, ca, ka);
    else exchange_one<Tile>(rc_a + (long)g * na * N, ga, na, N, u.sm, u.F, red, ca, ka);
    if (gb) {
      if (nb > 64) exchange_one_big<Tile>(rc_b + (long)g * nb * N, gb, nb, N, u.sm, u.F, F1, red, cb, kb);
      else exchange_one<Tile>(rc_b + (long)g * nb * N, gb, nb, N, u.sm, u.F, red, cb, kb);
    } else {
      cb = ca;
      kb = ka;
    }
    const cplx cs = cadd(ca, cb);
    const cplx c2 = cmul(cs, cs);
    sr += c2.re - ka.re - kb.re;
    si += c2.im - ka.im - kb.im;
  }
  if (threadIdx.x == 0) cstore(partial, w * gridDim.x + gblk, {0.5 * sr, 0.5 * si});
}

// E[w] = h0 + sum_s fac_s sum_{k,p} GT_s[w,k,p] rot_h1T_s[k,p] + sum_blocks partial[w, :]
__global__ __launch_bounds__(256) void energy_finalize_kernel(const double* __restrict__ gt_a,
                                                              const double* __restrict__ gt_b,
                                                              const double* __restrict__ h1t_a,
                                                              const double* __restrict__ h1t_b, int na, int nb, int N,
                                                              double h0, const double* __restrict__ partial, int nblk,
                                                              double* __restrict__ energy) {
  __shared__ double red[256];
  const long w = blockIdx.x;
  const int t = threadIdx.x;
  const double fac_a = gt_b ? 1.0 : 2.0;
  double er = 0, ei = 0;
  const double* ga = gt_a + 2 * w * (long)N * na;
  for (long e = t; e < (long)N * na; e += 256) {
    const cplx g = cload(ga, e);
    er += fac_a * g.re * h1t_a[e];
    ei += fac_a * g.im * h1t_a[e];
  }
  if (gt_b) {
    const double* gb = gt_b + 2 * w * (long)N * nb;
    for (long e = t; e < (long)N * nb; e += 256) {
      const cplx g = cload(gb, e);
      er += g.re * h1t_b[e];
      ei += g.im * h1t_b[e];
    }
  }
  er = block_sum<256>(er, red);
  ei = block_sum<256>(ei, red);
  if (t == 0) {
    double sr = h0 + er, si = ei;
    for (int b = 0; b < nblk; ++b) {
      const cplx p = cload(partial, w * nblk + b);
      sr += p.re;
      si += p.im;
    }
    cstore(energy, w, {sr, si});
  }
}

template <typename Tile>
static hipError_t launch_energy_t(const typename ExOps<Tile>::TA* rc_a, const typename ExOps<Tile>::TA* rc_b,
                                  const double* gt_a, const double* gt_b, const double* h1t_a, const double* h1t_b,
                                  int na, int nb, int N, int G, int W, double h0, int gb, double* partial,
                                  double* energy, hipStream_t s) {
  if (W <= 0) return hipSuccess;
  if (na > TROT_MAX_NOCC_BIG || nb > TROT_MAX_NOCC_BIG || na < 1 || (gt_b && nb < 1)) return hipErrorInvalidValue;
  const int nblk = (G + gb - 1) / gb;
  if (G > 0 && (na > 64 || (gt_b && nb > 64))) {
    hipLaunchKernelGGL(exchange_big_kernel<Tile>, dim3(nblk, W), dim3(256), 0, s, rc_a, rc_b, gt_a, gt_b, na, nb, N,
                       G, gb, partial);
  } else if (G > 0) {
    hipLaunchKernelGGL(exchange_kernel<Tile>, dim3(nblk, W), dim3(256), 0, s, rc_a, rc_b, gt_a, gt_b, na, nb, N, G,
                       gb, partial);
  }
  hipLaunchKernelGGL(energy_finalize_kernel, dim3(W), dim3(256), 0, s, gt_a, gt_b, h1t_a, h1t_b, na, nb, N, h0,
                     partial, G > 0 ? nblk : 0, energy);
  return hipGetLastError();
}

hipError_t launch_energy_finalize(const double* gt_a, const double* gt_b, const double* h1t_a, const double* h1t_b,
                                  int na, int nb, int N, double h0, const double* partial, int nblk, double* energy,
                                  int W, hipStream_t s) {
  if (W <= 0) return hipSuccess;
  hipLaunchKernelGGL(energy_finalize_kernel, dim3(W), dim3(256), 0, s, gt_a, gt_b, h1t_a, h1t_b, na, nb, N, h0,
                     partial, nblk, energy);
  return hipGetLastError();
}

hipError_t launch_energy(const double* rc_a, const double* rc_b, const double* gt_a, const double* gt_b,
                         const double* h1t_a, const double* h1t_b, int na, int nb, int N, int G, int W, double h0,
                         int gb, double* partial, double* energy, hipStream_t s) {
  return launch_energy_t<ExTile64>(rc_a, rc_b, gt_a, gt_b, h1t_a, h1t_b, na, nb, N, G, W, h0, gb, partial, energy, s);
}

hipError_t launch_energy_lp(const float* rc_a, const float* rc_b, const double* gt_a, const double* gt_b,
                            const double* h1t_a, const double* h1t_b, int na, int nb, int N, int G, int W, double h0,
                            int gb, double* partial, double* energy, int mode, hipStream_t s) {
  if (mode == MODE_F32)
    return launch_energy_t<ExTile32>(rc_a, rc_b, gt_a, gt_b, h1t_a, h1t_b, na, nb, N, G, W, h0, gb, partial, energy,
                                     s);
  if (mode == MODE_BF16X3)
    return launch_energy_t<ExTileB3>(rc_a, rc_b, gt_a, gt_b, h1t_a, h1t_b, na, nb, N, G, W, h0, gb, partial, energy,
                                     s);
  return hipErrorInvalidValue;
}
