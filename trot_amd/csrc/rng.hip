// Counter-based random numbers for the auxiliary fields and the SR comb offset.
//
// The reference draws its fields and the comb offset from JAX's threefry stream
// (trot/prop/afqmc.py:91-93, trot/prop/blocks.py:270-271), which cannot be reproduced without JAX.
// This stream is Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11), the generator numpy.random.Philox implements, so a CPU twin is checked
// bit for bit against numpy.
//
//   block j of draw c for purpose p:   x = Philox4x64-10(counter = (j, c, p, 0), key = (seed, 0))
//   normals (p = 0): Box-Muller on the 53-bit uniforms of (x0, x1) and (x2, x3):
//     u1 = ((x >> 11) + 1) 2^-53 in (0, 1],  u2 = (x >> 11) 2^-53 in [0, 1)
//     z = sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2)        -> elements 4 j .. 4 j + 3
//   uniform (p = 1): u = (x0 >> 11) 2^-53 of block 0
//
// (seed, c) come either by value or from a two-word device state {seed, counter}, so that a
// captured HIP graph draws fresh numbers on every replay; trot_rng_advance bumps the counter
// on the device.
#include "capi_util.hpp"
#include "common.hpp"

namespace {

__device__ __forceinline__ void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B97F4A7C15ull;
      k1 += 0xBB67AE8584CAA73Bull;
    }
    const uint64_t lo0 = 0xD2E7470EE14C6C93ull * c[0], hi0 = __umul64hi(0xD2E7470EE14C6C93ull, c[0]);
    const uint64_t lo1 = 0xCA5A826395121157ull * c[2], hi1 = __umul64hi(0xCA5A826395121157ull, c[2]);
    const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

__device__ __forceinline__ void seed_counter(const uint64_t* state, uint64_t seed, uint64_t ctr, uint64_t& s,
                                             uint64_t& c) {
  s = state ? state[0] : seed;
  c = state ? state[1] : ctr;
}

__global__ __launch_bounds__(256) void philox_raw_kernel(uint64_t seed, uint64_t ctr, uint64_t purpose, long nblk,
                                                         uint64_t* __restrict__ out) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (j >= nblk) return;
  uint64_t x[4] = {(uint64_t)j, ctr, purpose, 0};
  philox4x64_10(x, seed, 0);
#pragma unroll
  for (int e = 0; e < 4; ++e) out[4 * j + e] = x[e];
}

__device__ __forceinline__ void box_muller(uint64_t a, uint64_t b, double& z0, double& z1) {
  const double u1 = (double)((a >> 11) + 1) * 0x1.0p-53;
  const double u2 = (double)(b >> 11) * 0x1.0p-53;
  const double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincos(6.283185307179586 * u2, &s, &c);
  z0 = r * c;
  z1 = r * s;
}

__global__ __launch_bounds__(256) void philox_normal_kernel(const uint64_t* __restrict__ state, uint64_t seed,
                                                            uint64_t ctr, long n, double* __restrict__ out) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (4 * j >= n) return;
  uint64_t s, c;
  seed_counter(state, seed, ctr, s, c);
  uint64_t x[4] = {(uint64_t)j, c, 0, 0};
  philox4x64_10(x, s, 0);
  double z[4];
  box_muller(x[0], x[1], z[0], z[1]);
  box_muller(x[2], x[3], z[2], z[3]);
  if (4 * j + 4 <= n && (((uintptr_t)(out + 4 * j)) & 15) == 0) {
    reinterpret_cast<double2*>(out + 4 * j)[0] = make_double2(z[0], z[1]);
    reinterpret_cast<double2*>(out + 4 * j)[1] = make_double2(z[2], z[3]);
  } else {
    for (int e = 0; e < 4 && 4 * j + e < n; ++e) out[4 * j + e] = z[e];
  }
}

__global__ void philox_uniform_kernel(const uint64_t* __restrict__ state, uint64_t seed, uint64_t ctr,
                                      double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  uint64_t s, c;
  seed_counter(state, seed, ctr, s, c);
  uint64_t x[4] = {0, c, 1, 0};
  philox4x64_10(x, s, 0);
  out[0] = (double)(x[0] >> 11) * 0x1.0p-53;
}

__global__ void rng_advance_kernel(uint64_t* __restrict__ state, uint64_t by) {
  if (threadIdx.x == 0) state[1] += by;
}

}  // namespace

extern "C" {

int trot_philox_raw(uint64_t seed, uint64_t counter, uint64_t purpose, long n_blocks, uint64_t* out, void* stream) {
  if (!out || n_blocks < 0 || n_blocks > (1L << 31)) return fail(TROT_ERR_INVALID, "philox_raw: bad arguments");
  if (n_blocks == 0) return TROT_OK;
  hipLaunchKernelGGL(philox_raw_kernel, dim3((unsigned)((n_blocks + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, seed, counter, purpose, n_blocks, out);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_philox_normal(const uint64_t* state, uint64_t seed, uint64_t counter, long n, double* out, void* stream) {
  if (!out || n < 0 || (n + 3) / 4 > (1L << 31)) return fail(TROT_ERR_INVALID, "philox_normal: bad arguments");
  if (n == 0) return TROT_OK;
  const long nblk = (n + 3) / 4;
  hipLaunchKernelGGL(philox_normal_kernel, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     state, seed, counter, n, out);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_philox_uniform(const uint64_t* state, uint64_t seed, uint64_t counter, double* out, void* stream) {
  if (!out) return fail(TROT_ERR_INVALID, "philox_uniform: null output");
  hipLaunchKernelGGL(philox_uniform_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, state, seed, counter, out);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_rng_advance(uint64_t* state, uint64_t by, void* stream) {
  if (!state) return fail(TROT_ERR_INVALID, "rng_advance: null state");
  hipLaunchKernelGGL(rng_advance_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, state, by);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

}  // extern "C"
