// Shared definitions for the trot_amd HIP kernels (gfx950 / CDNA4 only).
//
// Complex data in HBM is interleaved (re, im) double pairs, the layout of
// torch.complex128 and of the reference's complex128 arrays.  Kernels
// de-interleave tiles into separate re / im planes in LDS and feed the
// planes to real MFMA instructions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

#define TROT_WAVE 64

struct cplx {
  double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }
__device__ __forceinline__ double cabs1(cplx a) { return fabs(a.re) + fabs(a.im); }
__device__ __forceinline__ cplx cinv(cplx a) {
  // Smith's algorithm for 1/a
  if (fabs(a.re) >= fabs(a.im)) {
    double r = a.im / a.re, d = a.re + a.im * r;
    return {1.0 / d, -r / d};
  } else {
    double r = a.re / a.im, d = a.re * r + a.im;
    return {r / d, -1.0 / d};
  }
}
__device__ __forceinline__ cplx cload(const double* p, long idx) {
  const double2 v = *reinterpret_cast<const double2*>(p + 2 * idx);
  return {v.x, v.y};
}
__device__ __forceinline__ void cstore(double* p, long idx, cplx v) {
  *reinterpret_cast<double2*>(p + 2 * idx) = make_double2(v.re, v.im);
}

// Symmetric N x N matrices in "sym-block" storage: only the 8 x 8 blocks
// (bi, bj) with bi <= bj, nb = ceil(N / 8) blocks per side, block-row major,
// each block row-major (64 elements); padding rows / columns hold zeros.
// p(bi, bj) = bi nb - bi (bi - 1) / 2 + (bj - bi).
__host__ __device__ __forceinline__ long symblk_index(int bi, int bj, int nb) {
  return (long)bi * nb - ((long)bi * (bi - 1)) / 2 + (bj - bi);
}
__host__ __device__ __forceinline__ long symblk_count(int nb) { return (long)nb * (nb + 1) / 2; }

// Element e (0 .. BM*BK-1) of an operand tile with rows m0.. and k k0.. of a
// sym-block matrix, enumerated block by block so that 64 consecutive e read
// one contiguous 64-element block.  Gives the tile-local (i, k) the element
// lands on (mirrored blocks are read transposed) and its storage offset.
template <int BK>
__device__ __forceinline__ bool symblk_coords(int e, int m0, int k0, int nb, int& i, int& k, long& off) {
  constexpr int KB = BK / 8;
  const int blk = e >> 6, wi = e & 63;
  const int rbl = blk / KB, kbl = blk - rbl * KB;
  const int rb = (m0 >> 3) + rbl, kb = (k0 >> 3) + kbl;
  const int a = wi >> 3, b = wi & 7;
  const bool up = rb <= kb;
  i = (rbl << 3) + (up ? a : b);
  k = (kbl << 3) + (up ? b : a);
  off = symblk_index(up ? rb : kb, up ? kb : rb, nb) * 64 + wi;
  return rb < nb && kb < nb;
}

// Deterministic block-wide sum of one double per thread (fixed tree).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}
