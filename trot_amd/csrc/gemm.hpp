// Batched real / complex GEMM on FP64 MFMA (v_mfma_f64_16x16x4_f64), gfx950.
//
//   C[b] = alpha * op(A[b]) * B[b]  (+ beta * C[b])            EPI_STORE
//   T[b] = alpha * A[b] * B[b];  O[b] += T[b]                  EPI_TAYLOR
//
// A and B are real or interleaved-complex; the product is complex whenever
// either operand is.  A complex product is four real MFMA streams into two
// accumulator planes (re, im); a real x complex product is two.
//
// Workgroup: 256 threads = 4 waves in a 2 x 2 grid; each wave owns a
// (BM/2) x (BN/2) block of 16 x 16 MFMA tiles.  K is staged through LDS in
// chunks of 16 with one chunk of global loads in flight in registers while
// the previous chunk is multiplied.  LDS tiles are stored k-major
// (As[plane][k][i], columns XOR-swizzled by k) so each MFMA operand fetch is 16 doubles of one row.
//
// MFMA f64 16x16x4 operand layout (lane l):  A[i = l % 16][k = l / 16],
// B[k = l / 16][j = l % 16];  D[i = 4 r + l / 16][j = l % 16] in register r (measured on
// gfx950; the f32 / bf16 16x16 variants put row 4 (l / 16) + r in register r instead).
#pragma once

#include "common.hpp"

enum { EPI_STORE = 0, EPI_TAYLOR = 1 };

struct GemmP {
  const double* A;
  long lda, sA;  // leading dim and batch stride, in elements (complex elements if complex)
  const double* B;
  long ldb, sB;
  double* C;
  long ldc, sC;
  double* C2;  // EPI_TAYLOR accumulation target
  long ldc2, sC2;
  int M, N, K;
  int Klast;      // K of the last batch when nonzero (split K with a shorter final slice)
  double ar, ai;  // alpha
  double beta;    // EPI_STORE: C = alpha * acc + beta * C (beta real)
};

// AKC: A[i][k] at i*lda + k (K contiguous), else at k*lda + i.
// BNC: B[k][j] at k*ldb + j (N contiguous), else at j*ldb + k.
// CONJA / CONJB: the complex operand enters conjugated.
// ASYM: A is a symmetric matrix in sym-block storage (common.hpp); lda = nb.
template <int BM, int BN, bool AC, bool BC, bool AKC, bool BNC, bool CONJA, bool ASYM = false, bool CONJB = false>
struct TileF64 {
  static constexpr int BK = 16;
  static constexpr int PA = AC ? 2 : 1;
  static constexpr int PB = BC ? 2 : 1;
  static constexpr int PC = (AC || BC) ? 2 : 1;
  static constexpr int WM = BM / 2, WN = BN / 2;
  static constexpr int TM = WM / 16, TN = WN / 16;
  static constexpr int PAD = 16;
  static constexpr int EA = BM * BK / 256;  // A elements per thread per chunk
  static constexpr int EB = BN * BK / 256;
  static_assert(EA >= 1 && EB >= 1, "tile too small");

  struct Smem {
    double As[PA][BK][BM + PAD];
    double Bs[PB][BK][BN + PAD];
  };
  // Column i of LDS row k is stored at i ^ (2 ((k >> 1) & 7)).  A staging store of K-contiguous global data
  // puts the 16 k of one i in 16 lanes: with the 80-double pitch alone rows k and k + 2 share banks (8-way
  // conflicts); the XOR gives the 32 doubles of a 32-lane pass 32 distinct bank pairs.  A fragment read
  // (16 consecutive i of one k per 16 lanes) is permuted inside aligned groups of 16 and stays
  // conflict-free.
  __device__ __forceinline__ static int sw(int k, int i) { return i ^ (((k >> 1) & 7) << 1); }
  typedef d4_t Acc[PC][TM][TN];

  // acc = op(A)[m0:m0+BM, :K] * B[:K, n0:n0+BN] (zero padded), A/B already offset to the batch.
  __device__ __forceinline__ static void run(const double* __restrict__ A, long lda, const double* __restrict__ B,
                                             long ldb, int M, int N, int K, int m0, int n0, Smem& sm, Acc& acc) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    double ra[EA][PA];
    double rb[EB][PB];
    int cia[EA];  // ASYM: tile-local (i << 8 | k) of each staged A element

    auto load_chunk = [&](int k0) {
#pragma unroll
      for (int r = 0; r < EA; ++r) {
        const int e = t + 256 * r;
        int i, k;
        long off;
        bool ok;
        if constexpr (ASYM) {
          ok = symblk_coords<BK>(e, m0, k0, (int)lda, i, k, off);
          cia[r] = (i << 8) | k;
        } else {
          if (AKC) { k = e % BK; i = e / BK; } else { i = e % BM; k = e / BM; }
          const int gi = m0 + i, gk = k0 + k;
          ok = gi < M && gk < K;
          off = AKC ? ((long)gi * lda + gk) : ((long)gk * lda + gi);
        }
        if (AC) {
          double2 v = ok ? *reinterpret_cast<const double2*>(A + 2 * off) : make_double2(0.0, 0.0);
          ra[r][0] = v.x;
          ra[r][PA - 1] = CONJA ? -v.y : v.y;
        } else {
          ra[r][0] = ok ? A[off] : 0.0;
        }
      }
#pragma unroll
      for (int r = 0; r < EB; ++r) {
        const int e = t + 256 * r;
        int j, k;
        if (BNC) { j = e % BN; k = e / BN; } else { k = e % BK; j = e / BK; }
        const int gj = n0 + j, gk = k0 + k;
        const bool ok = gj < N && gk < K;
        const long off = BNC ? ((long)gk * ldb + gj) : ((long)gj * ldb + gk);
        if (BC) {
          double2 v = ok ? *reinterpret_cast<const double2*>(B + 2 * off) : make_double2(0.0, 0.0);
          rb[r][0] = v.x;
          rb[r][PB - 1] = CONJB ? -v.y : v.y;
        } else {
          rb[r][0] = ok ? B[off] : 0.0;
        }
      }
    };
    auto store_chunk = [&]() {
#pragma unroll
      for (int r = 0; r < EA; ++r) {
        const int e = t + 256 * r;
        int i, k;
        if constexpr (ASYM) {
          i = cia[r] >> 8;
          k = cia[r] & 255;
        } else {
          if (AKC) { k = e % BK; i = e / BK; } else { i = e % BM; k = e / BM; }
        }
#pragma unroll
        for (int q = 0; q < PA; ++q) sm.As[q][k][sw(k, i)] = ra[r][q];
      }
#pragma unroll
      for (int r = 0; r < EB; ++r) {
        const int e = t + 256 * r;
        int j, k;
        if (BNC) { j = e % BN; k = e / BN; } else { k = e % BK; j = e / BK; }
#pragma unroll
        for (int q = 0; q < PB; ++q) sm.Bs[q][k][sw(k, j)] = rb[r][q];
      }
    };

#pragma unroll
    for (int q = 0; q < PC; ++q)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[q][i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

    const int nk = (K + BK - 1) / BK;
    load_chunk(0);
    for (int kc = 0; kc < nk; ++kc) {
      __syncthreads();
      store_chunk();
      __syncthreads();
      if (kc + 1 < nk) load_chunk((kc + 1) * BK);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4);
        double af[TM][PA], bf[TN][PB];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int q = 0; q < PA; ++q) af[i][q] = sm.As[q][kr][sw(kr, wm * WM + i * 16 + (lane & 15))];
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int q = 0; q < PB; ++q) bf[j][q] = sm.Bs[q][kr][sw(kr, wn * WN + j * 16 + (lane & 15))];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            if (AC && BC) {
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][0], bf[j][0], acc[0][i][j], 0, 0, 0);
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-af[i][1], bf[j][1], acc[0][i][j], 0, 0, 0);
              acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][0], bf[j][1], acc[1][i][j], 0, 0, 0);
              acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][1], bf[j][0], acc[1][i][j], 0, 0, 0);
            } else if (AC) {
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][0], bf[j][0], acc[0][i][j], 0, 0, 0);
              acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][1], bf[j][0], acc[1][i][j], 0, 0, 0);
            } else if (BC) {
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][0], bf[j][0], acc[0][i][j], 0, 0, 0);
              acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][0], bf[j][1], acc[1][i][j], 0, 0, 0);
            } else {
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][0], bf[j][0], acc[0][i][j], 0, 0, 0);
            }
          }
      }
    }
  }

  // (row, col) of register r of tile (i, j) for this lane, relative to (m0, n0)
  __device__ __forceinline__ static int row_of(int i, int r) {
    const int lane = threadIdx.x & 63, wm = (threadIdx.x >> 6) >> 1;
    return wm * WM + i * 16 + 4 * r + (lane >> 4);
  }
  __device__ __forceinline__ static int col_of(int j) {
    const int lane = threadIdx.x & 63, wn = (threadIdx.x >> 6) & 1;
    return wn * WN + j * 16 + (lane & 15);
  }
};

// One output tile (m0, n0) of batch b of problem p (nbatch batches): the whole work of a workgroup.
template <int BM, int BN, bool AC, bool BC, bool AKC, bool BNC, bool CONJA, int EPI, bool ASYM, bool CONJB = false>
__device__ __forceinline__ void gemm_f64_tile(const GemmP& p, int b, int nbatch, int m0, int n0,
                                              typename TileF64<BM, BN, AC, BC, AKC, BNC, CONJA, ASYM, CONJB>::Smem& sm) {
  typedef TileF64<BM, BN, AC, BC, AKC, BNC, CONJA, ASYM, CONJB> T;
  const double* A = p.A + (AC ? 2 : 1) * (long)b * p.sA;
  const double* B = p.B + (BC ? 2 : 1) * (long)b * p.sB;
  typename T::Acc acc;
  const int K = (p.Klast > 0 && b == nbatch - 1) ? p.Klast : p.K;
  T::run(A, p.lda, B, p.ldb, p.M, p.N, K, m0, n0, sm, acc);

  constexpr int PC = T::PC;
  double* C = p.C + PC * (long)b * p.sC;
  double* C2 = (EPI == EPI_TAYLOR) ? p.C2 + PC * (long)b * p.sC2 : nullptr;
#pragma unroll
  for (int i = 0; i < T::TM; ++i)
#pragma unroll
    for (int j = 0; j < T::TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + T::row_of(i, r);
        const int gj = n0 + T::col_of(j);
        if (gi >= p.M || gj >= p.N) continue;
        if (PC == 2) {
          const double xr = acc[0][i][j][r], xi = acc[1][i][j][r];
          double vr = p.ar * xr - p.ai * xi;
          double vi = p.ar * xi + p.ai * xr;
          const long off = (long)gi * p.ldc + gj;
          if (EPI == EPI_TAYLOR) {
            *reinterpret_cast<double2*>(C + 2 * off) = make_double2(vr, vi);
            const long off2 = (long)gi * p.ldc2 + gj;
            double2 o = *reinterpret_cast<double2*>(C2 + 2 * off2);
            *reinterpret_cast<double2*>(C2 + 2 * off2) = make_double2(o.x + vr, o.y + vi);
          } else {
            if (p.beta != 0.0) {
              double2 o = *reinterpret_cast<double2*>(C + 2 * off);
              vr += p.beta * o.x;
              vi += p.beta * o.y;
            }
            *reinterpret_cast<double2*>(C + 2 * off) = make_double2(vr, vi);
          }
        } else {
          double v = p.ar * acc[0][i][j][r];
          const long off = (long)gi * p.ldc + gj;
          if (p.beta != 0.0) v += p.beta * C[off];
          C[off] = v;
        }
      }
}

// Tile order: the (batch, row tile, column tile) triples in that order, column tile fastest, are dealt to the
// 8 XCDs in 8 contiguous ranges (workgroup w runs on XCD w % 8), so the tiles of one batch run on one XCD
// one after another and share the operand they have in common (a walker's right-hand matrix, a slice's
// panels) in that XCD's L2 instead of each XCD fetching it.
template <int BM, int BN, bool AC, bool BC, bool AKC, bool BNC, bool CONJA, int EPI, bool ASYM = false,
          bool CONJB = false>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmP p, int ntn, int ntm, int nbatch) {
  __shared__ typename TileF64<BM, BN, AC, BC, AKC, BNC, CONJA, ASYM, CONJB>::Smem sm;
  const long per_batch = (long)ntn * ntm, total = per_batch * nbatch;
  const long per = (total + 7) / 8;
  const long q = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (q >= total) return;
  const int b = (int)(q / per_batch);
  const int t = (int)(q - (long)b * per_batch);
  gemm_f64_tile<BM, BN, AC, BC, AKC, BNC, CONJA, EPI, ASYM, CONJB>(p, b, nbatch, (t / ntn) * BM, (t % ntn) * BN, sm);
}

// Many independent problems in one launch: problem i owns workgroups start[i] .. start[i+1] - 1, laid out
// (batch, row tile, column tile) with the column tile fastest, in that order (gemm_f64_kernel deals the same
// triples to the XCDs in ranges instead).  Each workgroup computes exactly what it would in a launch of its
// own problem.
template <int BM, int BN, bool AC, bool BC, bool AKC, bool BNC, bool CONJA, int EPI, bool ASYM = false>
__global__ __launch_bounds__(256) void gemm_f64_grouped_kernel(const GemmP* __restrict__ probs,
                                                               const int* __restrict__ start, int nprob) {
  __shared__ typename TileF64<BM, BN, AC, BC, AKC, BNC, CONJA, ASYM>::Smem sm;
  const int g = blockIdx.x;
  int lo = 0, hi = nprob;  // start[lo] <= g < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= g) lo = mid; else hi = mid;
  }
  const GemmP p = probs[lo];
  const int local = g - start[lo];
  const int ntn = (p.N + BN - 1) / BN, ntm = (p.M + BM - 1) / BM;
  const int tn = local % ntn, tm = (local / ntn) % ntm, b = local / (ntn * ntm);
  const int nbatch = (start[lo + 1] - start[lo]) / (ntn * ntm);  // the last batch may take Klast
  gemm_f64_tile<BM, BN, AC, BC, AKC, BNC, CONJA, EPI, ASYM>(p, b, nbatch, tm * BM, tn * BN, sm);
}

template <int BM, int BN>
inline int gemm_f64_workgroups(const GemmP& p, int batch) {
  return ((p.N + BN - 1) / BN) * ((p.M + BM - 1) / BM) * batch;
}

template <int BM, int BN, bool AC, bool BC, bool AKC, bool BNC, bool CONJA, int EPI, bool ASYM = false>
static hipError_t launch_gemm_f64_grouped(const GemmP* probs, const int* start, int nprob, int total_wg,
                                          hipStream_t s) {
  if (nprob <= 0 || total_wg <= 0) return hipSuccess;
  hipLaunchKernelGGL((gemm_f64_grouped_kernel<BM, BN, AC, BC, AKC, BNC, CONJA, EPI, ASYM>), dim3(total_wg), dim3(256),
                     0, s, probs, start, nprob);
  return hipGetLastError();
}

template <int BM, int BN, bool AC, bool BC, bool AKC, bool BNC, bool CONJA, int EPI, bool ASYM = false,
          bool CONJB = false>
static hipError_t launch_gemm_f64(const GemmP& p, int batch, hipStream_t s) {
  if (p.M <= 0 || p.N <= 0 || batch <= 0) return hipSuccess;
  const int ntn = (p.N + BN - 1) / BN, ntm = (p.M + BM - 1) / BM;
  const long total = (long)ntn * ntm * batch;
  const long nwg = (total + 7) / 8 * 8;
  if (nwg > 0x7fffffffL) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, AC, BC, AKC, BNC, CONJA, EPI, ASYM, CONJB>), dim3((unsigned)nwg),
                     dim3(256), 0, s, p, ntn, ntm, batch);
  return hipGetLastError();
}
