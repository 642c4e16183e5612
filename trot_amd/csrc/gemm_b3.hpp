// 3-pass split-BF16 "NT" GEMM on v_mfma_f32_16x16x32_bf16 (gfx950), FP32 accumulation:
//
//   C[m][n] = alpha * sum_k A[m][k] B[n][k]  (+ beta C)        A real or complex, B real
//
// Operands arrive pre-split into bf16 planes, x = hi + lo (hi = bf16(x),
// lo = bf16(x - hi)), each plane K-contiguous with rows padded to a multiple
// of 128 and K to a multiple of 32 (zeros), so the main loop has no bounds
// checks and every global load is a 16-byte vector:
//   A planes: [re_hi, re_lo, im_hi, im_lo] (complex) or [hi, lo], each (Mp, Kp)
//   B planes: [hi, lo], each (Np, Kp)
// a.b ~ lo_a.hi_b + hi_a.lo_b + hi_a.hi_b (the lo.lo term, ~2^-16 relative,
// is dropped).
//
// Workgroup 256 threads, tile 128 x 128 x 32; four waves in 2 x 2, each wave
// 64 x 64 = 4 x 4 MFMA tiles per plane: 96 MFMAs per K step for complex A.
// The next K chunk is loaded into registers while the current one is
// multiplied from LDS.  LDS rows hold 32 bf16 (64 B, four 16-byte segments)
// with segment s of row r at slot s ^ b3_swz(r) (see below), so neither the
// staging stores nor the fragment reads conflict on LDS banks.
//
// Tile order is XCD-aware: workgroup b runs on XCD b % 8, and XCD x is given
// a contiguous range of tiles ordered row-tile fastest, so consecutive
// workgroups on one XCD share a B column tile in that XCD's L2.
#pragma once

#include "common.hpp"

typedef unsigned short bf16raw;
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

struct B3P {
  const bf16raw* A;  // planes of (Mp, Kp)
  const bf16raw* B;  // planes of (Np, Kp)
  long a_plane, b_plane;  // plane strides (elements)
  int Kp;                 // padded K (multiple of 32); also the row stride
  int k_slice;            // K range per blockIdx.y (multiple of 32)
  int M, N;               // true sizes (output bounds)
  void* C;                // final output (when part == nullptr)
  long ldc;
  float* part;            // split K: raw FP32 partial sums [slice][M][N] (complex if AC), ld = N
  float ar, ai;
  float beta;
};

// XOR swizzle of the 16-byte segments of 64-byte LDS rows.  A fragment read takes one segment of
// 16 consecutive rows; a staging store takes the 4 segments of 4 consecutive rows (256 contiguous
// bytes, every bank once).  Rows r, r + 4, r + 8, r + 12 share the same 16 banks (64 banks of 4
// bytes); swz gives them slots 0, 2, 1, 3 in some order, so the 16 lanes of a fragment read that
// the LDS serves together touch 16 distinct 4-bank groups.
__device__ __forceinline__ int b3_swz(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }
// element offset of (row r, segment seg) in a tile of 64-byte rows
__device__ __forceinline__ int b3_lds(int r, int seg) { return r * 32 + 8 * (seg ^ b3_swz(r)); }

__device__ __forceinline__ void split_bf16_raw(float x, bf16raw& hi, bf16raw& lo) {
  const __bf16 h = (__bf16)x;
  const __bf16 l = (__bf16)(x - (float)h);
  hi = __builtin_bit_cast(bf16raw, h);
  lo = __builtin_bit_cast(bf16raw, l);
}

// Epilogue of a wave's 64 x 64 tile at (mrow, ncol): D[i = 4 (l / 16) + r][j = l % 16] in register r.
// slice: the K slice of a split-K product (its partial sums go to part[slice]); -1 takes blockIdx.y.
template <bool AC, typename TC>
__device__ __forceinline__ void b3_epilogue(const B3P& p, const f4_t (&acc)[AC ? 2 : 1][4][4], int mrow, int ncol,
                                            int lane, int slice = -1) {
  const long sl = slice < 0 ? (long)blockIdx.y : (long)slice;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = mrow + i * 16 + 4 * (lane >> 4) + r;
        const int gj = ncol + j * 16 + (lane & 15);
        if (gi >= p.M || gj >= p.N) continue;
        if (p.part) {  // split K: raw partial sums, reduced by b3_reduce_kernel
          const long po = (sl * p.M + gi) * p.N + gj;
          if constexpr (AC) {
            reinterpret_cast<float2*>(p.part)[po] = make_float2(acc[0][i][j][r], acc[1][i][j][r]);
          } else {
            p.part[po] = acc[0][i][j][r];
          }
          continue;
        }
        const long off = (long)gi * p.ldc + gj;
        if constexpr (AC) {
          const float xr = acc[0][i][j][r], xi = acc[1][i][j][r];
          const float vr = p.ar * xr - p.ai * xi, vi = p.ar * xi + p.ai * xr;
          if constexpr (__is_same(TC, unsigned)) {
            // split complex64: hi plane at C, lo plane M ldc words further; word = bf16 re | bf16 im << 16
            bf16raw rh, rl, ih, il;
            split_bf16_raw(vr, rh, rl);
            split_bf16_raw(vi, ih, il);
            unsigned* c = reinterpret_cast<unsigned*>(p.C);
            c[off] = (unsigned)rh | ((unsigned)ih << 16);
            c[(long)p.M * p.ldc + off] = (unsigned)rl | ((unsigned)il << 16);
          } else if constexpr (sizeof(TC) == 8) {
            double2* c = reinterpret_cast<double2*>(p.C) + off;
            double2 o = p.beta != 0.f ? *c : make_double2(0.0, 0.0);
            *c = make_double2((double)vr + (double)p.beta * o.x, (double)vi + (double)p.beta * o.y);
          } else {
            float2* c = reinterpret_cast<float2*>(p.C) + off;
            float2 o = p.beta != 0.f ? *c : make_float2(0.f, 0.f);
            *c = make_float2(vr + p.beta * o.x, vi + p.beta * o.y);
          }
        } else {
          const float v = p.ar * acc[0][i][j][r];
          if constexpr (sizeof(TC) == 8) {
            double* c = reinterpret_cast<double*>(p.C) + off;
            *c = (double)v + (p.beta != 0.f ? (double)p.beta * *c : 0.0);
          } else {
            float* c = reinterpret_cast<float*>(p.C) + off;
            *c = v + (p.beta != 0.f ? p.beta * *c : 0.f);
          }
        }
      }
}

template <bool AC, typename TC>
__global__ __launch_bounds__(256, 1) void gemm_b3_nt_kernel(B3P p, int n_mt, int n_nt) {
  constexpr int BM = 128, BN = 128, BK = 32;
  constexpr int PA = AC ? 4 : 2;  // A planes
  __shared__ __attribute__((aligned(16))) bf16raw As[PA][BM * BK];
  __shared__ __attribute__((aligned(16))) bf16raw Bs[2][BN * BK];

  // XCD-aware tile index
  const int T = n_mt * n_nt;
  const int per = (T + 7) / 8;
  const int b = blockIdx.x;
  const int tile = (b & 7) * per + (b >> 3);
  if (tile >= T) return;
  const int mt = tile % n_mt, nt = tile / n_mt;
  const int m0 = mt * BM, n0 = nt * BN;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  // staging: A PA planes x 2 segments, B 2 planes x 2 segments (16 B each), flat so that every
  // index is a compile-time constant after unrolling (the array then lives in VGPRs, not scratch)
  constexpr int NA = PA * 2, NB = 4;
  u32x4_t stg[NA + NB];
  const bf16raw* Ab = p.A + (long)m0 * p.Kp;
  const bf16raw* Bb = p.B + (long)n0 * p.Kp;
  const int row0 = t >> 2, kq = (t & 3) * 8;  // segment t; segment t + 256 is row0 + 64
  const int st0 = b3_lds(row0, t & 3), st1 = b3_lds(row0 + 64, t & 3);
#define B3_LOAD(k0)                                                                              \
  do {                                                                                           \
    const long off0 = (long)row0 * p.Kp + (k0) + kq;                                             \
    const long off1 = off0 + 64L * p.Kp;                                                         \
    _Pragma("unroll") for (int q = 0; q < PA; ++q) {                                             \
      stg[2 * q] = *reinterpret_cast<const u32x4_t*>(Ab + q * p.a_plane + off0);                   \
      stg[2 * q + 1] = *reinterpret_cast<const u32x4_t*>(Ab + q * p.a_plane + off1);               \
    }                                                                                            \
    _Pragma("unroll") for (int q = 0; q < 2; ++q) {                                              \
      stg[NA + 2 * q] = *reinterpret_cast<const u32x4_t*>(Bb + q * p.b_plane + off0);              \
      stg[NA + 2 * q + 1] = *reinterpret_cast<const u32x4_t*>(Bb + q * p.b_plane + off1);          \
    }                                                                                            \
  } while (0)
#define B3_STORE()                                                                               \
  do {                                                                                           \
    _Pragma("unroll") for (int q = 0; q < PA; ++q) {                                             \
      *reinterpret_cast<u32x4_t*>(&As[q][st0]) = stg[2 * q];                                       \
      *reinterpret_cast<u32x4_t*>(&As[q][st1]) = stg[2 * q + 1];                                   \
    }                                                                                            \
    _Pragma("unroll") for (int q = 0; q < 2; ++q) {                                              \
      *reinterpret_cast<u32x4_t*>(&Bs[q][st0]) = stg[NA + 2 * q];                                  \
      *reinterpret_cast<u32x4_t*>(&Bs[q][st1]) = stg[NA + 2 * q + 1];                              \
    }                                                                                            \
  } while (0)

  constexpr int PC = AC ? 2 : 1;
  f4_t acc[PC][4][4];
#pragma unroll
  for (int q = 0; q < PC; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[q][i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  const int kbeg = blockIdx.y * p.k_slice;
  const int nk = min(p.k_slice, p.Kp - kbeg) / BK;
  const int seg = lane >> 4;
  if (nk > 0) B3_LOAD(kbeg);
  for (int kc = 0; kc < nk; ++kc) {
    __syncthreads();
    B3_STORE();
    __syncthreads();
    if (kc + 1 < nk) B3_LOAD(kbeg + (kc + 1) * BK);
    bf16x8_t bh[4], bl[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = wn * 64 + j * 16 + (lane & 15);
      bh[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[0][b3_lds(col, seg)]);
      bl[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[1][b3_lds(col, seg)]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = wm * 64 + i * 16 + (lane & 15);
#pragma unroll
      for (int q = 0; q < PC; ++q) {
        const bf16x8_t ah = *reinterpret_cast<const bf16x8_t*>(&As[2 * q][b3_lds(row, seg)]);
        const bf16x8_t al = *reinterpret_cast<const bf16x8_t*>(&As[2 * q + 1][b3_lds(row, seg)]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          f4_t c = acc[q][i][j];
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], c, 0, 0, 0);
          acc[q][i][j] = c;
        }
      }
    }
  }

  b3_epilogue<AC, TC>(p, acc, m0 + wm * 64, n0 + wn * 64, lane);
}

// Wide form for complex A: workgroup 512 threads, tile 128 (complex rows) x 256 x 32, eight waves in
// 2 x 4 of 64 x 64.  Per K step the tile fetches 64 KB for 12.6 MFLOP (197 flop/B, against 131 for
// the 128 x 128 tile), which is what bounds the VHS and force-bias products when their operands
// stream from beyond L2.  LDS holds two stages (128 KB); one barrier per K step: the stage for
// k + 1 is stored while k is multiplied from the other, and the global loads for k + 2 are issued
// right after that store.  b_rows: padded row count of the B planes (rows beyond it read as zero).
constexpr int B3W_BM = 128, B3W_BN = 256, B3W_BK = 32;
constexpr int B3W_SMEM = 2 * (4 * B3W_BM + 2 * B3W_BN) * B3W_BK * 2;  // bytes

// Tile order: workgroup b runs on XCD b % 8 and XCD x takes a contiguous range of tiles; within it,
// strips of GM row tiles are walked column by column, row tile fastest, so that the workgroups
// resident on one XCD at a time cover about GM row tiles x (resident / GM) column tiles.
__device__ __forceinline__ bool b3_tile_of(int b, int n_mt, int n_nt, int GM, int& mt, int& nt) {
  const int T = n_mt * n_nt;
  const int per = (T + 7) / 8;
  const int tile = (b & 7) * per + (b >> 3);
  if (tile >= T) return false;
  const int strip = tile / (GM * n_nt);
  const int u = tile - strip * GM * n_nt;
  const int h = min(GM, n_mt - strip * GM);
  mt = strip * GM + u % h;
  nt = u / h;
  return true;
}

// Main loop of the wide form: acc[q][i][j] (q = re, im of A) for the wave's 64 x 64 sub-tile
// (rows wm * 64 .., columns wn * 64 ..; wm = wave / 4, wn = wave % 4) of the tile at (m0, n0).
//
// Operands go from global memory straight into LDS (global_load_lds_dwordx4: lane L of a wave writes the
// 16 bytes at LDS base + 16 L), without passing through registers.  A stage is 64 chunks of 1 KB (16 rows
// of 64 bytes: 32 of A over its 4 planes, 32 of B over its 2), 8 per wave.  Lane L of a chunk carries
// row L / 4 of the chunk into LDS slot L % 4, which the swizzle assigns to segment (L % 4) ^ b3_swz(row),
// so each lane picks the global address of that segment.  Two stages (128 KB): after the barrier that
// completes stage k, each wave reads all of its stage-k fragments, issues the loads of stage k + 1 into
// the other buffer (free: every wave is past its reads of stage k - 1) and runs the MFMAs of stage k,
// which cover those loads.  B rows at or beyond b_rows are read from row b_rows - 1: output column j
// depends on B row j only, and columns beyond the true count are never written.  Ends on a barrier
// with no load in flight, so the caller may reuse smem.
// TOUT: the MFMAs take B as their first operand, so acc[q][i][j][r] holds row 16 i + (lane & 15), column
// 16 j + 4 (lane >> 4) + r of the wave's sub-tile (the same products in the same order; b3w_epilogue);
// otherwise row 16 i + 4 (lane >> 4) + r, column 16 j + (lane & 15) (b3_epilogue's layout).
template <bool TOUT = false>
__device__ __forceinline__ void b3w_mainloop(const B3P& p, int m0, int n0, int b_rows, char* smem,
                                             f4_t (&acc)[2][4][4], int slice) {
  constexpr int BM = B3W_BM, BN = B3W_BN, BK = B3W_BK;
  constexpr int ASTAGE = 4 * BM * BK, BSTAGE = 2 * BN * BK;  // bf16 elements per stage
  typedef __attribute__((address_space(3))) bf16raw lds_t;
  lds_t* As = (lds_t*)smem;                  // [2 stages][4 planes][BM * BK]
  lds_t* Bs = As + 2 * ASTAGE;               // [2 stages][2 planes][BN * BK]

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  const int crow = lane >> 2;  // row of the lane within its chunk
  const bf16raw* asrc[4];
  const bf16raw* bsrc[4];
  int aoff[4], boff[4];  // wave-uniform LDS element offsets of the chunks within a stage
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int c = 4 * wave + i;
    const int qa = c >> 3, ra = 16 * (c & 7) + crow;
    asrc[i] = p.A + qa * p.a_plane + (long)(m0 + ra) * p.Kp + 8 * ((lane & 3) ^ b3_swz(ra));
    aoff[i] = qa * BM * BK + 16 * (c & 7) * BK;
    const int qb = c >> 4, rb = 16 * (c & 15) + crow;
    const int gb = min(n0 + rb, b_rows - 1);
    bsrc[i] = p.B + qb * p.b_plane + (long)gb * p.Kp + 8 * ((lane & 3) ^ b3_swz(rb));
    boff[i] = qb * BN * BK + 16 * (c & 15) * BK;
  }
  auto issue = [&](int buf, int k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      __builtin_amdgcn_global_load_lds((const void*)(asrc[i] + k0), (__attribute__((address_space(3))) void*)(As + buf * ASTAGE + aoff[i]),
                                       16, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      __builtin_amdgcn_global_load_lds((const void*)(bsrc[i] + k0), (__attribute__((address_space(3))) void*)(Bs + buf * BSTAGE + boff[i]),
                                       16, 0, 0);
  };

#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[q][i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  const int kbeg = slice * p.k_slice;
  const int nk = min(p.k_slice, p.Kp - kbeg) / BK;
  const int seg = lane >> 4;
  if (nk > 0) issue(0, kbeg);
  for (int kc = 0; kc < nk; ++kc) {
    const int cur = kc & 1;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's loads of stage kc have landed
    __syncthreads();                                   // and every wave's
    const lds_t* A = As + cur * ASTAGE;
    const lds_t* B = Bs + cur * BSTAGE;
    // B fragments and row tile 0's A fragments first; the loads of the next stage; then the MFMAs of row
    // tile i while the A fragments of row tile i + 1 are read
    bf16x8_t bh[4], bl[4], ah[2][2], al[2][2];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = wn * 64 + j * 16 + (lane & 15);
      bh[j] = *reinterpret_cast<const __attribute__((address_space(3))) bf16x8_t*>(B + b3_lds(col, seg));
      bl[j] = *reinterpret_cast<const __attribute__((address_space(3))) bf16x8_t*>(B + BN * BK + b3_lds(col, seg));
    }
    auto read_a = [&](int i) {
      const int row = wm * 64 + i * 16 + (lane & 15);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        ah[i & 1][q] = *reinterpret_cast<const __attribute__((address_space(3))) bf16x8_t*>(A + 2 * q * BM * BK + b3_lds(row, seg));
        al[i & 1][q] = *reinterpret_cast<const __attribute__((address_space(3))) bf16x8_t*>(A + (2 * q + 1) * BM * BK + b3_lds(row, seg));
      }
    };
    read_a(0);
    if (kc + 1 < nk) issue(cur ^ 1, kbeg + (kc + 1) * BK);
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i < 3) read_a(i + 1);
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          f4_t c = acc[q][i][j];
          if constexpr (TOUT) {
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i & 1][q], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i & 1][q], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i & 1][q], c, 0, 0, 0);
          } else {
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i & 1][q], bh[j], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i & 1][q], bl[j], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i & 1][q], bh[j], c, 0, 0, 0);
          }
          acc[q][i][j] = c;
        }
      // Issue order of row tile i: one of the next row tile's four fragment reads after each group of 4
      // MFMAs, then the last 8 MFMAs.  Left to itself the compiler clusters the reads and waits for them a
      // few MFMAs later (its LDS waits here drain every outstanding read); spread out, each read has 4-16
      // MFMAs of cover before the wait at the next row tile.
      if (i < 3) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);  // MFMA
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
      } else {
        __builtin_amdgcn_sched_group_barrier(0x008, 24, 0);
      }
    }
    __builtin_amdgcn_s_setprio(0);
  }
  __syncthreads();
}

// Epilogue of the wide kernels for the TOUT layout (complex A): the four values of a lane are consecutive
// columns gj0 .. gj0 + 3 of row gi, so each plane of a run leaves as one 16-byte store (two for complex64
// and FP32 partials) when the output's leading dimension keeps runs 16-byte aligned; otherwise, and for
// runs cut by the column bound, element by element.  Each element is computed by the same expressions as
// b3_epilogue's.
template <typename TC>
__device__ __forceinline__ void b3w_epilogue(const B3P& p, const f4_t (&acc)[2][4][4], int mrow, int ncol, int lane,
                                             int slice = -1) {
  const long sl = slice < 0 ? (long)blockIdx.y : (long)slice;
  const bool c16 = (reinterpret_cast<uintptr_t>(p.C) & 15) == 0;
  bool vec;
  if (p.part) {
    vec = (p.N & 1) == 0;
  } else if constexpr (__is_same(TC, unsigned)) {
    vec = c16 && (p.ldc & 3) == 0;
  } else if constexpr (sizeof(TC) == 4) {
    vec = c16 && (p.ldc & 1) == 0 && p.beta == 0.f;
  } else {
    vec = false;  // complex128: one 16-byte store per element either way
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gi = mrow + i * 16 + (lane & 15);
      const int gj0 = ncol + j * 16 + 4 * (lane >> 4);
      if (gi >= p.M) continue;
      float vr[4], vi[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float xr = acc[0][i][j][r], xi = acc[1][i][j][r];
        vr[r] = p.part ? xr : p.ar * xr - p.ai * xi;
        vi[r] = p.part ? xi : p.ar * xi + p.ai * xr;
      }
      if (vec && gj0 + 3 < p.N) {
        if (p.part) {
          float4* o = reinterpret_cast<float4*>(reinterpret_cast<float2*>(p.part) + (sl * p.M + gi) * p.N + gj0);
          o[0] = make_float4(vr[0], vi[0], vr[1], vi[1]);
          o[1] = make_float4(vr[2], vi[2], vr[3], vi[3]);
        } else if constexpr (__is_same(TC, unsigned)) {
          u32x4_t hw, lw;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            bf16raw rh, rl, ih, il;
            split_bf16_raw(vr[r], rh, rl);
            split_bf16_raw(vi[r], ih, il);
            hw[r] = (unsigned)rh | ((unsigned)ih << 16);
            lw[r] = (unsigned)rl | ((unsigned)il << 16);
          }
          unsigned* c = reinterpret_cast<unsigned*>(p.C) + (long)gi * p.ldc + gj0;
          *reinterpret_cast<u32x4_t*>(c) = hw;
          *reinterpret_cast<u32x4_t*>(c + (long)p.M * p.ldc) = lw;
        } else if constexpr (sizeof(TC) == 4) {
          float4* o = reinterpret_cast<float4*>(reinterpret_cast<float2*>(p.C) + (long)gi * p.ldc + gj0);
          o[0] = make_float4(vr[0], vi[0], vr[1], vi[1]);
          o[1] = make_float4(vr[2], vi[2], vr[3], vi[3]);
        }
        continue;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = gj0 + r;
        if (gj >= p.N) continue;
        if (p.part) {
          reinterpret_cast<float2*>(p.part)[(sl * p.M + gi) * p.N + gj] = make_float2(vr[r], vi[r]);
          continue;
        }
        const long off = (long)gi * p.ldc + gj;
        if constexpr (__is_same(TC, unsigned)) {
          bf16raw rh, rl, ih, il;
          split_bf16_raw(vr[r], rh, rl);
          split_bf16_raw(vi[r], ih, il);
          unsigned* c = reinterpret_cast<unsigned*>(p.C);
          c[off] = (unsigned)rh | ((unsigned)ih << 16);
          c[(long)p.M * p.ldc + off] = (unsigned)rl | ((unsigned)il << 16);
        } else if constexpr (sizeof(TC) == 8) {
          double2* c = reinterpret_cast<double2*>(p.C) + off;
          double2 o = p.beta != 0.f ? *c : make_double2(0.0, 0.0);
          *c = make_double2((double)vr[r] + (double)p.beta * o.x, (double)vi[r] + (double)p.beta * o.y);
        } else {
          float2* c = reinterpret_cast<float2*>(p.C) + off;
          float2 o = p.beta != 0.f ? *c : make_float2(0.f, 0.f);
          *c = make_float2(vr[r] + p.beta * o.x, vi[r] + p.beta * o.y);
        }
      }
    }
}

template <typename TC>
__global__ __launch_bounds__(512, 1) void gemm_b3w_kernel(B3P p, int n_mt, int n_nt, int b_rows) {
  __shared__ __attribute__((aligned(16))) char smem[B3W_SMEM];
  int mt, nt;
  if (!b3_tile_of(blockIdx.x, n_mt, n_nt, 8, mt, nt)) return;
  const int m0 = mt * B3W_BM, n0 = nt * B3W_BN;
  f4_t acc[2][4][4];
  b3w_mainloop<true>(p, m0, n0, b_rows, smem, acc, blockIdx.y);
  const int wave = threadIdx.x >> 6;
  b3w_epilogue<TC>(p, acc, m0 + (wave >> 2) * 64, n0 + (wave & 3) * 64, threadIdx.x & 63);
}

// The wide GEMMs of many problems in one launch (the batching form's force bias and VHS in mixed precision):
// problem i owns workgroups start[i] .. start[i+1] - 1, S slices of its tiles_pad (a multiple of 8) tile
// workgroups, and runs exactly the workgroup body of its own launch (same tile order, same K slices).
struct B3GProb {
  B3P p;
  int n_mt, n_nt, b_rows, S, tiles_pad, c_kind;
  double ar, ai, beta;  // FP64 alpha and beta of the slice reduction (split K)
};

template <typename TC>
__global__ __launch_bounds__(512, 1) void gemm_b3w_grouped_kernel(const B3GProb* __restrict__ probs,
                                                                  const int* __restrict__ start, int nprob) {
  __shared__ __attribute__((aligned(16))) char smem[B3W_SMEM];
  int lo = 0, hi = nprob;  // start[lo] <= g < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= (int)blockIdx.x) lo = mid; else hi = mid;
  }
  const B3GProb q = probs[lo];
  const int local = blockIdx.x - start[lo];
  const int slice = local / q.tiles_pad, tb = local - slice * q.tiles_pad;
  int mt, nt;
  if (!b3_tile_of(tb, q.n_mt, q.n_nt, 8, mt, nt)) return;
  const int m0 = mt * B3W_BM, n0 = nt * B3W_BN;
  f4_t acc[2][4][4];
  b3w_mainloop<true>(q.p, m0, n0, q.b_rows, smem, acc, slice);
  const int wave = threadIdx.x >> 6;
  b3w_epilogue<TC>(q.p, acc, m0 + (wave >> 2) * 64, n0 + (wave & 3) * 64, threadIdx.x & 63, slice);
}

// C (complex128, M x N, ldc) = alpha * sum_s part[s] (+ beta C), slices summed in order in FP64.
__device__ __forceinline__ void b3_reduce_c128_body(const float* __restrict__ part, int S, int M, int N,
                                                    double* __restrict__ C, long ldc, double ar, double ai, double beta,
                                                    long e) {
  if (e >= (long)M * N) return;
  const int i = (int)(e / N), j = (int)(e - (e / N) * N);
  double sr = 0.0, si = 0.0;
  for (int s = 0; s < S; ++s) {
    const float2 v = reinterpret_cast<const float2*>(part)[(long)s * M * N + e];
    sr += v.x;
    si += v.y;
  }
  double2* c = reinterpret_cast<double2*>(C) + (long)i * ldc + j;
  const double2 o = beta != 0.0 ? *c : make_double2(0.0, 0.0);
  *c = make_double2(ar * sr - ai * si + beta * o.x, ar * si + ai * sr + beta * o.y);
}

static __global__ void b3_reduce_c128_kernel(const float* __restrict__ part, int S, int M, int N, double* __restrict__ C,
                                             long ldc, double ar, double ai, double beta) {
  b3_reduce_c128_body(part, S, M, N, C, ldc, ar, ai, beta, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Split a row-major FP64 matrix (rows x K, leading dim ld; complex if CX) into bf16 planes of
// (rows_pad, Kp), zero padded.  Planes: [re_hi, re_lo, im_hi, im_lo] or [hi, lo].
// TRANS: element (r, k) of the source sits at k * ld + r (column-major source).
template <bool CX, bool TRANS, typename TS>
__device__ __forceinline__ void split_planes_body(const TS* __restrict__ src, long ld, int rows, int K, int rows_pad,
                                                  int Kp, bf16raw* __restrict__ dst, long e) {
  const long total = (long)rows_pad * Kp;
  if (e >= total) return;
  const int r = (int)(e / Kp), k = (int)(e - (e / Kp) * Kp);
  const bool ok = r < rows && k < K;
  float vr = 0.f, vi = 0.f;
  if (ok) {
    const long so = TRANS ? (long)k * ld + r : (long)r * ld + k;
    if (CX) {
      vr = (float)src[2 * so];
      vi = (float)src[2 * so + 1];
    } else {
      vr = (float)src[so];
    }
  }
  bf16raw h, l;
  split_bf16_raw(vr, h, l);
  dst[e] = h;
  dst[total + e] = l;
  if (CX) {
    split_bf16_raw(vi, h, l);
    dst[2 * total + e] = h;
    dst[3 * total + e] = l;
  }
}

template <bool CX, bool TRANS, typename TS>
__global__ void split_planes_kernel(const TS* __restrict__ src, long ld, int rows, int K, int rows_pad, int Kp,
                                    bf16raw* __restrict__ dst) {
  split_planes_body<CX, TRANS, TS>(src, ld, rows, K, rows_pad, Kp, dst, (long)blockIdx.x * blockDim.x + threadIdx.x);
}
