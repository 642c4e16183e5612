// Reduced-precision batched real / complex GEMM tiles on gfx950 MFMA.
//
//   MODE_F32     v_mfma_f32_16x16x4_f32, FP32 operands, FP32 accumulation
//   MODE_BF16X3  v_mfma_f32_16x16x32_bf16 three times per K block:
//                x = hi + lo with hi = bf16(x), lo = bf16(x - hi);
//                a.b ~ hi_a.hi_b + hi_a.lo_b + lo_a.hi_b, FP32 accumulation
//
// Operands may live in HBM as double or float (real or interleaved complex);
// they are converted (and split, for BF16X3) on their way into LDS, so the
// FP64 walker-side operands need no separate cast pass.  Complex products are
// four real MFMA streams into two accumulator planes, as in the FP64 tile.
//
// Workgroup: 256 threads, 4 waves in 2 x 2, each wave (BM/2) x (BN/2).
// MFMA f32 16x16 operand layouts (lane l): A[i = l % 16][k = l / 16] (f32) or
// A[i = l % 16][k = 8 (l / 16) + e] (bf16, e = 0..7); B likewise with j for i;
// D[i = 4 (l / 16) + r][j = l % 16] in register r.
#pragma once

#include "common.hpp"

enum { MODE_F32 = 1, MODE_BF16X3 = 2 };

template <typename T>
struct Ld;
template <>
struct Ld<double> {
  __device__ __forceinline__ static float re(const double* p, long off) { return (float)p[off]; }
  __device__ __forceinline__ static float2 cx(const double* p, long off) {
    const double2 v = *reinterpret_cast<const double2*>(p + 2 * off);
    return make_float2((float)v.x, (float)v.y);
  }
};
template <>
struct Ld<float> {
  __device__ __forceinline__ static float re(const float* p, long off) { return p[off]; }
  __device__ __forceinline__ static float2 cx(const float* p, long off) {
    return *reinterpret_cast<const float2*>(p + 2 * off);
  }
};

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// ASYM: A is a symmetric matrix in sym-block storage (common.hpp); lda = nb.
template <int BM, int BN, int MODE, bool AC, bool BC, typename TA, typename TB, bool AKC, bool BNC, bool CONJA,
          bool ASYM = false>
struct TileLP {
  static constexpr int BK = MODE == MODE_F32 ? 16 : 32;
  static constexpr int PA = AC ? 2 : 1;
  static constexpr int PB = BC ? 2 : 1;
  static constexpr int PC = (AC || BC) ? 2 : 1;
  static constexpr int WM = BM / 2, WN = BN / 2;
  static constexpr int TM = WM / 16, TN = WN / 16;
  static constexpr int EA = BM * BK / 256;
  static constexpr int EB = BN * BK / 256;
  static constexpr int PAD = 16;  // f32 k-major rows
  static constexpr int KP = 8;    // bf16 i-major rows (keeps 16-byte alignment)
  static_assert(EA >= 1 && EB >= 1, "tile too small");

  struct SmemF32 {
    float As[PA][BK][BM + PAD];
    float Bs[PB][BK][BN + PAD];
  };
  // MODE_F32: column i of LDS row k is stored at i ^ (4 ((k >> 2) & 3)).  A staging store of K-contiguous
  // global data puts the 16 k of one i in 16 lanes, and with the 80-float pitch alone rows k, k + 4, k + 8,
  // k + 12 share banks (4-way conflicts); with the XOR the 64 lanes of a store take 64 distinct banks.  A
  // fragment read (16 consecutive i of one k per 16 lanes) is permuted inside aligned groups of 16.
  __device__ __forceinline__ static int sw(int k, int i) { return i ^ (((k >> 2) & 3) << 2); }
  struct SmemBF16 {
    __bf16 Ah[PA][BM][BK + KP];
    __bf16 Al[PA][BM][BK + KP];
    __bf16 Bh[PB][BN][BK + KP];
    __bf16 Bl[PB][BN][BK + KP];
  };
  typedef typename std::conditional<MODE == MODE_F32, SmemF32, SmemBF16>::type Smem;
  typedef f4_t Acc[PC][TM][TN];

  __device__ __forceinline__ static void run(const TA* __restrict__ A, long lda, const TB* __restrict__ B, long ldb,
                                             int M, int N, int K, int m0, int n0, Smem& sm, Acc& acc) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    float ra[EA][PA];
    float rb[EB][PB];
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
          const float2 v = ok ? Ld<TA>::cx(A, off) : make_float2(0.f, 0.f);
          ra[r][0] = v.x;
          ra[r][PA - 1] = CONJA ? -v.y : v.y;
        } else {
          ra[r][0] = ok ? Ld<TA>::re(A, off) : 0.f;
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
          const float2 v = ok ? Ld<TB>::cx(B, off) : make_float2(0.f, 0.f);
          rb[r][0] = v.x;
          rb[r][PB - 1] = v.y;
        } else {
          rb[r][0] = ok ? Ld<TB>::re(B, off) : 0.f;
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
        for (int q = 0; q < PA; ++q) {
          if constexpr (MODE == MODE_F32) {
            sm.As[q][k][sw(k, i)] = ra[r][q];
          } else {
            __bf16 h, l;
            split_bf16(ra[r][q], h, l);
            sm.Ah[q][i][k] = h;
            sm.Al[q][i][k] = l;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < EB; ++r) {
        const int e = t + 256 * r;
        int j, k;
        if (BNC) { j = e % BN; k = e / BN; } else { k = e % BK; j = e / BK; }
#pragma unroll
        for (int q = 0; q < PB; ++q) {
          if constexpr (MODE == MODE_F32) {
            sm.Bs[q][k][sw(k, j)] = rb[r][q];
          } else {
            __bf16 h, l;
            split_bf16(rb[r][q], h, l);
            sm.Bh[q][j][k] = h;
            sm.Bl[q][j][k] = l;
          }
        }
      }
    };

#pragma unroll
    for (int q = 0; q < PC; ++q)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[q][i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

    const int nk = (K + BK - 1) / BK;
    load_chunk(0);
    for (int kc = 0; kc < nk; ++kc) {
      __syncthreads();
      store_chunk();
      __syncthreads();
      if (kc + 1 < nk) load_chunk((kc + 1) * BK);
      if constexpr (MODE == MODE_F32) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
          const int kr = kk + (lane >> 4);
          float af[TM][PA], bf[TN][PB];
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
                acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][0], bf[j][0], acc[0][i][j], 0, 0, 0);
                acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-af[i][1], bf[j][1], acc[0][i][j], 0, 0, 0);
                acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][0], bf[j][1], acc[1][i][j], 0, 0, 0);
                acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][1], bf[j][0], acc[1][i][j], 0, 0, 0);
              } else if (AC) {
                acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][0], bf[j][0], acc[0][i][j], 0, 0, 0);
                acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][1], bf[j][0], acc[1][i][j], 0, 0, 0);
              } else if (BC) {
                acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][0], bf[j][0], acc[0][i][j], 0, 0, 0);
                acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][0], bf[j][1], acc[1][i][j], 0, 0, 0);
              } else {
                acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][0], bf[j][0], acc[0][i][j], 0, 0, 0);
              }
            }
        }
      } else {
        const int ko = 8 * (lane >> 4);
        bf16x8_t ah[TM][PA], al[TM][PA], bh[TN][PB], bl[TN][PB];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int q = 0; q < PA; ++q) {
            const int row = wm * WM + i * 16 + (lane & 15);
            ah[i][q] = *reinterpret_cast<const bf16x8_t*>(&sm.Ah[q][row][ko]);
            al[i][q] = *reinterpret_cast<const bf16x8_t*>(&sm.Al[q][row][ko]);
          }
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int q = 0; q < PB; ++q) {
            const int col = wn * WN + j * 16 + (lane & 15);
            bh[j][q] = *reinterpret_cast<const bf16x8_t*>(&sm.Bh[q][col][ko]);
            bl[j][q] = *reinterpret_cast<const bf16x8_t*>(&sm.Bl[q][col][ko]);
          }
        auto mma3 = [&](const bf16x8_t& xh, const bf16x8_t& xl, const bf16x8_t& yh, const bf16x8_t& yl, f4_t c) {
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, yh, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, yl, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, yh, c, 0, 0, 0);
          return c;
        };
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            if (AC && BC) {
              // imaginary part of A enters the real plane negated: negate its bf16 parts
              bf16x8_t nh, nl;
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                nh[e] = -ah[i][1][e];
                nl[e] = -al[i][1][e];
              }
              acc[0][i][j] = mma3(ah[i][0], al[i][0], bh[j][0], bl[j][0], acc[0][i][j]);
              acc[0][i][j] = mma3(nh, nl, bh[j][1], bl[j][1], acc[0][i][j]);
              acc[1][i][j] = mma3(ah[i][0], al[i][0], bh[j][1], bl[j][1], acc[1][i][j]);
              acc[1][i][j] = mma3(ah[i][1], al[i][1], bh[j][0], bl[j][0], acc[1][i][j]);
            } else if (AC) {
              acc[0][i][j] = mma3(ah[i][0], al[i][0], bh[j][0], bl[j][0], acc[0][i][j]);
              acc[1][i][j] = mma3(ah[i][1], al[i][1], bh[j][0], bl[j][0], acc[1][i][j]);
            } else if (BC) {
              acc[0][i][j] = mma3(ah[i][0], al[i][0], bh[j][0], bl[j][0], acc[0][i][j]);
              acc[1][i][j] = mma3(ah[i][0], al[i][0], bh[j][1], bl[j][1], acc[1][i][j]);
            } else {
              acc[0][i][j] = mma3(ah[i][0], al[i][0], bh[j][0], bl[j][0], acc[0][i][j]);
            }
          }
      }
    }
  }

  __device__ __forceinline__ static int row_of(int i, int r) {
    const int lane = threadIdx.x & 63, wm = (threadIdx.x >> 6) >> 1;
    return wm * WM + i * 16 + 4 * (lane >> 4) + r;
  }
  __device__ __forceinline__ static int col_of(int j) {
    const int lane = threadIdx.x & 63, wn = (threadIdx.x >> 6) & 1;
    return wn * WN + j * 16 + (lane & 15);
  }
};

// Output writers for the low-precision GEMM kernel.
template <typename T>
struct St;
template <>
struct St<double> {
  __device__ __forceinline__ static void cx(double* p, long off, float vr, float vi, double beta) {
    double2* q = reinterpret_cast<double2*>(p + 2 * off);
    double2 o = beta != 0.0 ? *q : make_double2(0.0, 0.0);
    *q = make_double2((double)vr + beta * o.x, (double)vi + beta * o.y);
  }
  __device__ __forceinline__ static void re(double* p, long off, float v, double beta) {
    p[off] = (double)v + (beta != 0.0 ? beta * p[off] : 0.0);
  }
};
template <>
struct St<float> {
  __device__ __forceinline__ static void cx(float* p, long off, float vr, float vi, double beta) {
    float2* q = reinterpret_cast<float2*>(p + 2 * off);
    float2 o = beta != 0.0 ? *q : make_float2(0.f, 0.f);
    *q = make_float2(vr + (float)beta * o.x, vi + (float)beta * o.y);
  }
  __device__ __forceinline__ static void re(float* p, long off, float v, double beta) {
    p[off] = v + (beta != 0.0 ? (float)beta * p[off] : 0.f);
  }
};

template <typename TA, typename TB, typename TC>
struct GemmLPP {
  const TA* A;
  long lda, sA;
  const TB* B;
  long ldb, sB;
  TC* C;
  long ldc, sC;
  double* C2;  // EPI_TAYLOR: FP64 complex accumulation target
  long ldc2, sC2;
  int M, N, K;
  float ar, ai;
  double beta;
};

// EPI 0: C = alpha acc (+ beta C).  EPI 1 (Taylor): C = alpha acc; C2 (FP64) += C.
template <int BM, int BN, int MODE, bool AC, bool BC, typename TA, typename TB, typename TC, bool AKC, bool BNC,
          bool CONJA, int EPI, bool ASYM = false>
__global__ __launch_bounds__(256) void gemm_lp_kernel(GemmLPP<TA, TB, TC> p) {
  typedef TileLP<BM, BN, MODE, AC, BC, TA, TB, AKC, BNC, CONJA, ASYM> T;
  __shared__ typename T::Smem sm;
  const int b = blockIdx.z;
  const int m0 = blockIdx.y * BM;
  const int n0 = blockIdx.x * BN;
  const TA* A = p.A + (AC ? 2 : 1) * (long)b * p.sA;
  const TB* B = p.B + (BC ? 2 : 1) * (long)b * p.sB;
  typename T::Acc acc;
  T::run(A, p.lda, B, p.ldb, p.M, p.N, p.K, m0, n0, sm, acc);
  constexpr int PC = T::PC;
  TC* C = p.C + PC * (long)b * p.sC;
  double* C2 = EPI == 1 ? p.C2 + 2 * (long)b * p.sC2 : nullptr;
#pragma unroll
  for (int i = 0; i < T::TM; ++i)
#pragma unroll
    for (int j = 0; j < T::TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + T::row_of(i, r);
        const int gj = n0 + T::col_of(j);
        if (gi >= p.M || gj >= p.N) continue;
        const long off = (long)gi * p.ldc + gj;
        if (PC == 2) {
          const float xr = acc[0][i][j][r], xi = acc[1][i][j][r];
          const float vr = p.ar * xr - p.ai * xi;
          const float vi = p.ar * xi + p.ai * xr;
          St<TC>::cx(C, off, vr, vi, EPI == 1 ? 0.0 : p.beta);
          if (EPI == 1) {
            double2* q = reinterpret_cast<double2*>(C2 + 2 * ((long)gi * p.ldc2 + gj));
            const double2 o = *q;
            *q = make_double2(o.x + (double)vr, o.y + (double)vi);
          }
        } else {
          St<TC>::re(C, off, p.ar * acc[0][i][j][r], p.beta);
        }
      }
}

template <int BM, int BN, int MODE, bool AC, bool BC, typename TA, typename TB, typename TC, bool AKC, bool BNC,
          bool CONJA, int EPI, bool ASYM = false>
static hipError_t launch_gemm_lp(const GemmLPP<TA, TB, TC>& p, int batch, hipStream_t s) {
  if (p.M <= 0 || p.N <= 0 || batch <= 0) return hipSuccess;
  dim3 grid((p.N + BN - 1) / BN, (p.M + BM - 1) / BM, batch);
  if (grid.y > 65535 || grid.z > 65535) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((gemm_lp_kernel<BM, BN, MODE, AC, BC, TA, TB, TC, AKC, BNC, CONJA, EPI, ASYM>), grid, dim3(256), 0, s,
                     p);
  return hipGetLastError();
}
