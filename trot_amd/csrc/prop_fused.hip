// The whole FP64 propagation of small walker sets in one grouped launch (the batching form's fragments):
//
//   w <- E . T_K(i sqrt(dt) V) . E w,   T_K(aV) u = u + sum_k T_k,  T_k = (a / k) V T_{k-1}
//
// (trot/prop/chol_afqmc_ops.py:130-131, 152-162, 198-222; trot/prop/utils.py:20-29).  One workgroup of 8 waves
// per walker and spin set: the one-body product, the six Taylor terms and the last one-body product run with the
// current right operand (N x n complex, zero padded to multiples of 16) resident in LDS, so a walker's series
// needs no launch per term and reads its V straight from the sym-block storage the VHS GEMM wrote.  Each wave
// owns 16-row tiles of the output (tile r: rows 16 (wave + 8 r)) and all 16-column tiles.
//
// Bits: every output element is accumulated by exactly the MFMA sequence of the tile GEMM of gemm.hpp that
// trot_one_body_apply and trot_taylor_apply_symblk run (v_mfma_f64_16x16x4_f64 over K in ascending steps of 4,
// K zero padded to a multiple of 16, re += ar br then re += (-ai) bi, im += ar bi then im += ai br), with the same
// epilogue expressions, so the results equal those of the per-walker-set launches bit for bit.
#include "capi_util.hpp"
#include "common.hpp"
#include "kernels.hpp"

struct FusedProb {
  const double* vsym;  // V, sym-block storage, (W, vcols) complex
  long vcols;
  const double* expm;  // E = exp(-dt/2 h1_eff), (N, N) real
  double* w;           // walkers (W, N, n) complex, in / out
  double* sum;         // the Taylor running sum, (W, N, n) complex scratch
  int W, N, n, nb, Kp, ncp, nct;
  double sqrt_dt;
  int n_terms;
};

namespace {

constexpr int FT = 512;  // threads per workgroup
constexpr size_t FUSED_LDS_MAX = 160 * 1024;

__device__ __forceinline__ cplx vsym_at(const double* __restrict__ v, int i, int k, int nb) {
  const int bi = i >> 3, bk = k >> 3;
  if (bi >= nb || bk >= nb) return {0.0, 0.0};
  const long off = bi <= bk ? symblk_index(bi, bk, nb) * 64 + (i & 7) * 8 + (k & 7)
                            : symblk_index(bk, bi, nb) * 64 + (k & 7) * 8 + (i & 7);
  return cload(v, off);
}

// acc[r][c] = A[rows of tile r, :] . B[:, 16 c ..] over K = Kp, A = V (complex, VA) or E (real).  Two K steps
// per iteration with all their operands (V or E from global memory, the term from LDS) loaded before the
// first MFMA, so the loads of the second step overlap the first step's MFMAs: 7 % faster than one step per
// iteration; four steps need more than the 256 registers a wave has at 8 waves per workgroup.
template <bool VA>
__device__ __forceinline__ void fused_stage(const FusedProb& q, const double* __restrict__ V, const double* Bre,
                                            const double* Bim, d4_t (&acc)[2][3][2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrt = q.Kp >> 4;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[r][c][0] = acc[r][c][1] = d4_t{0.0, 0.0, 0.0, 0.0};
  const int kl = lane >> 4, il = lane & 15;
  // two K steps per iteration (Kp is a multiple of 16): every operand of both is loaded before the first MFMA
  for (int k0 = 0; k0 < q.Kp; k0 += 8) {
    double ar[2][2], ai[2][2], br[2][3], bi[2][3];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = k0 + 4 * h + kl;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int rt = wave + 8 * r;
        const int i = 16 * rt + il;
        const bool ok = rt < nrt && i < q.N && k < q.N;
        if (VA) {
          const cplx a = ok ? vsym_at(V, i, k, q.nb) : cplx{0.0, 0.0};
          ar[h][r] = a.re;
          ai[h][r] = a.im;
        } else {
          ar[h][r] = ok ? q.expm[(long)i * q.N + k] : 0.0;
          ai[h][r] = 0.0;
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int e = k * q.ncp + 16 * c + il;
        br[h][c] = c < q.nct ? Bre[e] : 0.0;
        bi[h][c] = c < q.nct ? Bim[e] : 0.0;
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (c >= q.nct) break;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          if (wave + 8 * r >= nrt) break;
          if (VA) {
            acc[r][c][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[h][r], br[h][c], acc[r][c][0], 0, 0, 0);
            acc[r][c][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[h][r], bi[h][c], acc[r][c][0], 0, 0, 0);
            acc[r][c][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[h][r], bi[h][c], acc[r][c][1], 0, 0, 0);
            acc[r][c][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[h][r], br[h][c], acc[r][c][1], 0, 0, 0);
          } else {
            acc[r][c][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[h][r], br[h][c], acc[r][c][0], 0, 0, 0);
            acc[r][c][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[h][r], bi[h][c], acc[r][c][1], 0, 0, 0);
          }
        }
      }
  }
}

// The wave's output values v = alpha acc (alpha = ar + i ai), as the tile GEMM's epilogue forms them, into LDS
// (all padded positions) and, for valid rows and columns, into `out` (store) or `addto` (+=).
__device__ __forceinline__ void fused_epilogue(const FusedProb& q, const d4_t (&acc)[2][3][2], double par, double pai,
                                               double* Bre, double* Bim, bool to_lds, double* out, double* addto) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrt = q.Kp >> 4;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int rt = wave + 8 * r;
    if (rt >= nrt) break;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (c >= q.nct) break;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int gi = 16 * rt + 4 * e + (lane >> 4);
        const int gj = 16 * c + (lane & 15);
        const double xr = acc[r][c][0][e], xi = acc[r][c][1][e];
        const double vr = par * xr - pai * xi;
        const double vi = par * xi + pai * xr;
        if (to_lds) {
          Bre[gi * q.ncp + gj] = vr;
          Bim[gi * q.ncp + gj] = vi;
        }
        if (gi < q.N && gj < q.n) {
          const long o = (long)gi * q.n + gj;
          if (out) *reinterpret_cast<double2*>(out + 2 * o) = make_double2(vr, vi);
          if (addto) {
            const double2 s = *reinterpret_cast<const double2*>(addto + 2 * o);
            *reinterpret_cast<double2*>(addto + 2 * o) = make_double2(s.x + vr, s.y + vi);
          }
        }
      }
    }
  }
}

__device__ __forceinline__ void fused_load(const FusedProb& q, const double* __restrict__ src, double* Bre,
                                           double* Bim) {
  for (int e = threadIdx.x; e < q.Kp * q.ncp; e += FT) {
    const int k = e / q.ncp, j = e - (e / q.ncp) * q.ncp;
    const cplx x = (k < q.N && j < q.n) ? cload(src, (long)k * q.n + j) : cplx{0.0, 0.0};
    Bre[e] = x.re;
    Bim[e] = x.im;
  }
}

__global__ __launch_bounds__(FT) void prop_fused_grouped_kernel(const FusedProb* __restrict__ probs,
                                                               const int* __restrict__ start, int nprob) {
  extern __shared__ double sh[];
  const int g = blockIdx.x;
  int lo = 0, hi = nprob;  // start[lo] <= g < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= g) lo = mid; else hi = mid;
  }
  const FusedProb q = probs[lo];
  const long wk = g - start[lo];
  double* Bre = sh;
  double* Bim = sh + (long)q.Kp * q.ncp;
  double* w = q.w + 2 * wk * (long)q.N * q.n;
  double* sum = q.sum + 2 * wk * (long)q.N * q.n;
  const double* V = q.vsym + 2 * wk * q.vcols;
  d4_t acc[2][3][2];

  fused_load(q, w, Bre, Bim);
  __syncthreads();
  fused_stage<false>(q, nullptr, Bre, Bim, acc);  // u = E w: the running sum starts at u (trot_one_body_apply)
  __syncthreads();
  fused_epilogue(q, acc, 1.0, 0.0, Bre, Bim, true, sum, nullptr);
  __syncthreads();
  for (int k = 1; k <= q.n_terms; ++k) {  // T_k = (i sqrt(dt) / k) V T_{k-1}; sum += T_k (EPI_TAYLOR)
    fused_stage<true>(q, V, Bre, Bim, acc);
    __syncthreads();
    fused_epilogue(q, acc, 0.0, q.sqrt_dt / k, Bre, Bim, k < q.n_terms, nullptr, sum);
    __syncthreads();
  }
  fused_load(q, sum, Bre, Bim);
  __syncthreads();
  fused_stage<false>(q, nullptr, Bre, Bim, acc);  // w = E sum
  fused_epilogue(q, acc, 1.0, 0.0, Bre, Bim, false, w, nullptr);
}

inline int pad16(int x) { return (x + 15) / 16 * 16; }

}  // namespace

extern "C" {

size_t trot_fused_prop_problem_bytes(void) { return sizeof(FusedProb); }

long trot_fused_prop_problem(const double* vsym, const double* expm, double* w, double* sum, int W, int N, int n,
                             double sqrt_dt, int n_terms, void* desc) {
  if (!vsym || !expm || !w || !sum || !desc || w == sum) return fail(TROT_ERR_INVALID, "fused_prop_problem: bad pointers");
  if (W < 1 || N < 1 || n < 1 || n > N || n_terms < 0) return fail(TROT_ERR_INVALID, "fused_prop_problem: bad sizes");
  const int Kp = pad16(N), ncp = pad16(n);
  if ((size_t)Kp * ncp * 16 > FUSED_LDS_MAX || Kp > 16 * 16 || ncp > 48)
    return fail(TROT_ERR_UNSUPPORTED, "fused_prop_problem: walker set too large for the LDS-resident series");
  FusedProb p{};
  p.vsym = vsym;
  p.nb = (N + 7) / 8;
  p.vcols = symblk_count(p.nb) * 64;
  p.expm = expm;
  p.w = w;
  p.sum = sum;
  p.W = W;
  p.N = N;
  p.n = n;
  p.Kp = Kp;
  p.ncp = ncp;
  p.nct = ncp / 16;
  p.sqrt_dt = sqrt_dt;
  p.n_terms = n_terms;
  *reinterpret_cast<FusedProb*>(desc) = p;
  return W;
}

// probs: device array of nprob descriptors; start: device prefix sums of their walker counts (nprob + 1 ints);
// max_lds: the largest Kp ncp 16 over the problems (bytes of dynamic LDS).
int trot_fused_prop_grouped(const void* probs, const int* start, int nprob, int total_wg, long max_lds,
                            void* stream) {
  if (!probs || !start || nprob < 1 || total_wg < 1 || max_lds < 1 || (size_t)max_lds > FUSED_LDS_MAX)
    return fail(TROT_ERR_INVALID, "fused_prop_grouped: bad arguments");
  if ((size_t)max_lds > 65536)
    HIPCHK(hipFuncSetAttribute((const void*)prop_fused_grouped_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)max_lds));
  hipLaunchKernelGGL(prop_fused_grouped_kernel, dim3(total_wg), dim3(FT), (size_t)max_lds, (hipStream_t)stream,
                     (const FusedProb*)probs, start, nprob);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

}  // extern "C"
