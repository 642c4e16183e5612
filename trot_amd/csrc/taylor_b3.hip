// Taylor series of exp(i sqrt(dt) V) w on 3-pass split-BF16 MFMA (trot/prop/utils.py:20-29).
//
// Per walker: V (N x N, complex, symmetric), w (N x n, complex128).
//   T_0 = w;  T_k = (i sqrt(dt) / k) V T_{k-1};  w += T_k        k = 1 .. K
//
// Operand storage (bf16 planes x = hi + lo, [re_hi, re_lo, im_hi, im_lo]):
//   V planes  (W, 4, Np, Np), Np = N rounded up to 128, zero padded: row i, K = column, contiguous;
//   T planes  (W, 4, np, Np), np = n rounded up to 64: T^T, so that K (= orbital) is contiguous.
// V planes are expanded from the sym-block complex64 VHS once per step; T_0 planes are the split
// transpose of w; every Taylor kernel writes the next term's planes and adds the term to w (FP64).
//
// taylor_b3_kernel: grid (Np/128 * np/64, W), 256 threads, tile 128 (orbitals) x 64 (occupied) x 32,
// waves 2 x 2 of 64 x 32 (4 x 2 MFMA tiles).  Complex x complex = 4 real products x 3 passes:
// re += Ar.Br + Ai.(-Bi), im += Ar.Bi + Ai.Br (the -Bi fragment is a sign flip of its bf16 words).
#include "capi_util.hpp"
#include "common.hpp"
#include "gemm_b3.hpp"

__device__ __forceinline__ bf16x8_t neg_bf16x8(bf16x8_t x) {
  u32x4_t u = __builtin_bit_cast(u32x4_t, x);
  u ^= (u32x4_t){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  return __builtin_bit_cast(bf16x8_t, u);
}

__global__ __launch_bounds__(256, 1) void taylor_b3_kernel(const bf16raw* __restrict__ vp,
                                                           const bf16raw* __restrict__ tp, bf16raw* __restrict__ tn,
                                                           double* __restrict__ w, int N, int n, int Np, int np_,
                                                           float c) {
  constexpr int BM = 128, BN = 64, BK = 32, LDR = 40;
  __shared__ __attribute__((aligned(16))) bf16raw As[4][BM][LDR];
  __shared__ __attribute__((aligned(16))) bf16raw Bs[4][BN][LDR];
  const int n_rt = Np / BM;
  const int rt = blockIdx.x % n_rt, ct = blockIdx.x / n_rt;
  const long b = blockIdx.y;
  const int m0 = rt * BM, n0 = ct * BN;
  const long vplane = (long)Np * Np, tplane = (long)np_ * Np;
  const bf16raw* Ab = vp + b * 4 * vplane + (long)m0 * Np;
  const bf16raw* Bb = tp + b * 4 * tplane + (long)n0 * Np;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int row0 = t >> 2, kq = (t & 3) * 8;  // A: segments t, t + 256 (rows row0, row0 + 64); B: segment t

  u32x4_t stg[12];
#define TB3_LOAD(k0)                                                                              \
  do {                                                                                            \
    const long off0 = (long)row0 * Np + (k0) + kq;                                                \
    const long off1 = off0 + 64L * Np;                                                            \
    _Pragma("unroll") for (int q = 0; q < 4; ++q) {                                               \
      stg[2 * q] = *reinterpret_cast<const u32x4_t*>(Ab + q * vplane + off0);                     \
      stg[2 * q + 1] = *reinterpret_cast<const u32x4_t*>(Ab + q * vplane + off1);                 \
      stg[8 + q] = *reinterpret_cast<const u32x4_t*>(Bb + q * tplane + off0);                     \
    }                                                                                             \
  } while (0)
#define TB3_STORE()                                                                               \
  do {                                                                                            \
    _Pragma("unroll") for (int q = 0; q < 4; ++q) {                                               \
      *reinterpret_cast<u32x4_t*>(&As[q][row0][kq]) = stg[2 * q];                                 \
      *reinterpret_cast<u32x4_t*>(&As[q][row0 + 64][kq]) = stg[2 * q + 1];                        \
      *reinterpret_cast<u32x4_t*>(&Bs[q][row0][kq]) = stg[8 + q];                                 \
    }                                                                                             \
  } while (0)

  f4_t acc[2][4][2];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[q][i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  const int nk = Np / BK;
  const int ko = 8 * (lane >> 4);
  TB3_LOAD(0);
  for (int kc = 0; kc < nk; ++kc) {
    __syncthreads();
    TB3_STORE();
    __syncthreads();
    if (kc + 1 < nk) TB3_LOAD((kc + 1) * BK);
    // B fragments: [j][plane] with plane 0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo; im negated copies
    bf16x8_t bf[2][4], bn[2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = wn * 32 + j * 16 + (lane & 15);
#pragma unroll
      for (int q = 0; q < 4; ++q) bf[j][q] = *reinterpret_cast<const bf16x8_t*>(&Bs[q][col][ko]);
      bn[j][0] = neg_bf16x8(bf[j][2]);
      bn[j][1] = neg_bf16x8(bf[j][3]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = wm * 64 + i * 16 + (lane & 15);
      bf16x8_t a[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) a[q] = *reinterpret_cast<const bf16x8_t*>(&As[q][row][ko]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        f4_t cr = acc[0][i][j], ci = acc[1][i][j];
        // re += Ar.Br - Ai.Bi
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bf[j][0], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][1], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][0], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[3], bn[j][0], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], bn[j][1], cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], bn[j][0], cr, 0, 0, 0);
        // im += Ar.Bi + Ai.Br
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bf[j][2], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][3], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][2], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[3], bf[j][0], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], bf[j][1], ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], bf[j][0], ci, 0, 0, 0);
        acc[0][i][j] = cr;
        acc[1][i][j] = ci;
      }
    }
  }
#undef TB3_LOAD
#undef TB3_STORE

  // epilogue: T = i c acc -> (re, im) = (-c acc_im, c acc_re).  Lane holds rows 4 (l/16) + r, column l % 16.
  bf16raw* Tn = tn + b * 4 * tplane;
  double* wb = w + 2 * b * (long)N * n;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gi0 = m0 + wm * 64 + i * 16 + 4 * (lane >> 4);
      const int gp = n0 + wn * 32 + j * 16 + (lane & 15);
      bf16raw hr[4], lr[4], hi[4], li[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float tre = -c * acc[1][i][j][r], tim = c * acc[0][i][j][r];
        split_bf16_raw(tre, hr[r], lr[r]);
        split_bf16_raw(tim, hi[r], li[r]);
        const int gi = gi0 + r;
        if (gi < N && gp < n) {
          double2* q = reinterpret_cast<double2*>(wb) + (long)gi * n + gp;
          const double2 o = *q;
          *q = make_double2(o.x + (double)tre, o.y + (double)tim);
        }
      }
      // 4 consecutive rows of T^T row gp: one 8-byte store per plane
      const long o = (long)gp * Np + gi0;
      typedef unsigned short us4 __attribute__((ext_vector_type(4)));
      *reinterpret_cast<us4*>(Tn + 0 * tplane + o) = (us4){hr[0], hr[1], hr[2], hr[3]};
      *reinterpret_cast<us4*>(Tn + 1 * tplane + o) = (us4){lr[0], lr[1], lr[2], lr[3]};
      *reinterpret_cast<us4*>(Tn + 2 * tplane + o) = (us4){hi[0], hi[1], hi[2], hi[3]};
      *reinterpret_cast<us4*>(Tn + 3 * tplane + o) = (us4){li[0], li[1], li[2], li[3]};
    }
}

// V planes (W, 4, Np, Np) from the sym-block complex64 VHS (W, P*64): one 8 x 8 block (bi <= bj)
// per 64-thread workgroup, written at (bi, bj) and, transposed, at (bj, bi).  Grid (nb, nb, W).
__global__ __launch_bounds__(64) void expand_vhs_b3_kernel(const float* __restrict__ vsym, bf16raw* __restrict__ vp,
                                                           int nb, int Np, long cols) {
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  __shared__ float2 blk[64];
  const long b = blockIdx.z;
  const int t = threadIdx.x;
  blk[t] = reinterpret_cast<const float2*>(vsym)[b * cols + symblk_index(bi, bj, nb) * 64 + t];
  __syncthreads();
  const long plane = (long)Np * Np;
  bf16raw* base = vp + b * 4 * plane;
  const int a = t >> 3, e = t & 7;
  {  // block (bi, bj): element (a, e) at row 8 bi + a, column 8 bj + e
    const float2 v = blk[t];
    const long o = (long)(8 * bi + a) * Np + 8 * bj + e;
    bf16raw h, l;
    split_bf16_raw(v.x, h, l);
    base[o] = h;
    base[plane + o] = l;
    split_bf16_raw(v.y, h, l);
    base[2 * plane + o] = h;
    base[3 * plane + o] = l;
  }
  if (bi != bj) {  // mirrored block (bj, bi): element (a, e) is V[8 bi + e][8 bj + a]
    const float2 v = blk[e * 8 + a];
    const long o = (long)(8 * bj + a) * Np + 8 * bi + e;
    bf16raw h, l;
    split_bf16_raw(v.x, h, l);
    base[o] = h;
    base[plane + o] = l;
    split_bf16_raw(v.y, h, l);
    base[2 * plane + o] = h;
    base[3 * plane + o] = l;
  }
}

// T^T planes (W, 4, np, Np) from walkers w (W, N, n) complex128, zero padded.  Grid (Np/64, np/64, W),
// 256 threads; a 64 x 64 tile goes through LDS so that reads run along p and writes along i.
__global__ __launch_bounds__(256) void split_transpose_w_kernel(const double* __restrict__ w, bf16raw* __restrict__ tp,
                                                                int N, int n, int Np, int np_) {
  __shared__ double2 tile[64][65];
  const int i0 = blockIdx.x * 64, p0 = blockIdx.y * 64;
  const long b = blockIdx.z;
  const int t = threadIdx.x;
  const double2* wb = reinterpret_cast<const double2*>(w) + b * (long)N * n;
  for (int e = t; e < 64 * 64; e += 256) {
    const int ii = e >> 6, pp = e & 63;  // read along p
    const int gi = i0 + ii, gp = p0 + pp;
    tile[ii][pp] = (gi < N && gp < n) ? wb[(long)gi * n + gp] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const long plane = (long)np_ * Np;
  bf16raw* base = tp + b * 4 * plane;
  for (int e = t; e < 64 * 64; e += 256) {
    const int pp = e >> 6, ii = e & 63;  // write along i
    const double2 v = tile[ii][pp];
    const long o = (long)(p0 + pp) * Np + i0 + ii;
    bf16raw h, l;
    split_bf16_raw((float)v.x, h, l);
    base[o] = h;
    base[plane + o] = l;
    split_bf16_raw((float)v.y, h, l);
    base[2 * plane + o] = h;
    base[3 * plane + o] = l;
  }
}

static inline int pad_up(int x, int m) { return (x + m - 1) / m * m; }

extern "C" {

size_t trot_taylor_b3_workspace(int W, int N, int n) {
  const size_t Np = pad_up(N, 128), np_ = pad_up(n, 64);
  return align256((size_t)W * 4 * Np * Np * 2) + 2 * align256((size_t)W * 4 * np_ * Np * 2);
}

// w (complex128, in/out) <- sum_{k=0..K} (i sqrt(dt) V)^k / k! w with V given in sym-block storage as
// complex64 (W, trot_symblk_cols(N)).  All products on split-BF16 MFMA, the sum in FP64.
int trot_taylor_apply_b3_symblk(const float* vhs_symblk, double* w, int W, int N, int n, double sqrt_dt, int n_terms,
                                void* ws, size_t ws_bytes, void* stream) {
  if (!vhs_symblk || !w) return fail(TROT_ERR_INVALID, "taylor_apply_b3: null pointer");
  if (W < 1 || N < 1 || n < 1 || n_terms < 0 || W > 65535) return fail(TROT_ERR_INVALID, "taylor_apply_b3: bad sizes");
  if (ws_bytes < trot_taylor_b3_workspace(W, N, n)) return fail(TROT_ERR_WORKSPACE, "taylor_apply_b3: workspace");
  const int Np = pad_up(N, 128), np_ = pad_up(n, 64), nb = (N + 7) / 8;
  Carve cv{(char*)ws, ws_bytes, 0};
  bf16raw* vp = cv.take<bf16raw>((size_t)W * 4 * Np * Np);
  bf16raw* t0 = cv.take<bf16raw>((size_t)W * 4 * np_ * Np);
  bf16raw* t1 = cv.take<bf16raw>((size_t)W * 4 * np_ * Np);
  hipStream_t s = (hipStream_t)stream;
  if (n_terms == 0) return TROT_OK;
  // padding rows / columns of V beyond 8 nb are never written by the expansion: clear them
  if (Np > 8 * nb) HIPCHK(hipMemsetAsync(vp, 0, (size_t)W * 4 * Np * Np * 2, s));
  hipLaunchKernelGGL(expand_vhs_b3_kernel, dim3(nb, nb, W), dim3(64), 0, s, vhs_symblk, vp, nb, Np,
                     (long)symblk_count(nb) * 64);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(split_transpose_w_kernel, dim3(Np / 64, np_ / 64, W), dim3(256), 0, s, w, t0, N, n, Np, np_);
  HIPCHK(hipGetLastError());
  bf16raw* prev = t0;
  bf16raw* next = t1;
  for (int k = 1; k <= n_terms; ++k) {
    hipLaunchKernelGGL(taylor_b3_kernel, dim3((Np / 128) * (np_ / 64), W), dim3(256), 0, s, vp, prev, next, w, N, n,
                       Np, np_, (float)(sqrt_dt / k));
    HIPCHK(hipGetLastError());
    bf16raw* tmp = prev;
    prev = next;
    next = tmp;
  }
  return TROT_OK;
}

}  // extern "C"
