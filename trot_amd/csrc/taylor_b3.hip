// Walker propagation  w <- E . T_K(i sqrt(dt) V) . E w  on 3-pass split-BF16 MFMA
// (trot/prop/chol_afqmc_ops.py:130-131, 152-162; trot/prop/utils.py:20-29).
//
// Per walker: V (N x N, complex, symmetric, sym-block complex64 as the VHS GEMM writes it),
// w (N x n, complex128), E = exp(-dt/2 h1_eff) (N x N, real, shared by all walkers).
//
//   one body:  w <- w + (E - I) w              identity term exact in FP64
//   Taylor:    T_0 = w;  T_k = (i sqrt(dt) / k) V T_{k-1};  w <- w + sum_k T_k
//
// The right operand of every product is a set of split planes of the walker-side matrix
// (x = hi + lo, bf16, [re_hi, re_lo, im_hi, im_lo]) stored transposed, (W, 4, np, Np) with
// Np = N rounded up to 128 and np = n rounded up to 64, zero padded, so that K (the orbital)
// is contiguous.  Each kernel writes the planes its successor reads: T_k for the next term,
// the updated walker after the last term and after the first one-body product.
// The Taylor terms k >= 1 are summed in FP32 (S, (W, Np, np) complex64); the last term adds
// S + T_K to w in FP64.  The terms are small next to w, so FP32 rounding of their sum stays
// far below the split-BF16 product error.
//
// prop_b3_kernel: one 128 (orbitals) x 64 (occupied) output tile per workgroup; 4 waves of 64 x 32
// (K staged 32 at a time) or 8 waves of 32 x 32 (K staged 64 at a time).  The V tile is read straight from
// sym-block storage: the 8 x 8 blocks above the diagonal land in LDS as stored, the mirrored
// ones transposed, and both are split into bf16 planes on the way.  Workgroups are ordered so
// that the row tiles of one walker run on one XCD (workgroup b runs on XCD b % 8) and share
// its T planes and the mirrored V blocks in that XCD's L2.
// Complex x complex = 4 real products x 3 passes: re += Ar.Br + Ai.(-Bi), im += Ar.Bi + Ai.Br
// (the -Bi fragment is a sign flip of its bf16 words); real E x complex = 2 x 3 passes.
#include <vector>

#include "capi_util.hpp"
#include "common.hpp"
#include "gemm_b3.hpp"

namespace {

constexpr int PB_BM = 128, PB_BN = 64;

// Waves per workgroup: 8 (32 x 32 each, four per SIMD at two workgroups per CU) for the one-body product
// and the one-pass term, whose registers allow it; 4 (64 x 32 each) for the three-pass term, which needs
// more than 128 registers a lane and would drop to one workgroup per CU with 8.  The MFMA sequence of each
// output element is the same either way.
// The second launch bound (waves per SIMD) keeps two workgroups per CU: 4 for 8 waves, 2 for 4.
template <bool REAL_A, int VM>
struct PropWaves {
  static constexpr int value = (REAL_A || VM == 2) ? 8 : 4;
};

struct PropB3P {
  const float* vsym;    // Taylor: V, sym-block complex64 (W, vcols)
  const unsigned* vh;   // Taylor, split V: hi plane (W, vcols) of words bf16 re | bf16 im << 16
  const unsigned* vl;   //   lo plane, same layout
  long vcols;           // complex numbers per walker in vsym
  const bf16raw* ep;    // one body: (E - I) planes [hi, lo], each (Np, Np)
  const bf16raw* tp;    // input planes (W, 4, np, Np)
  bf16raw* tn;          // output planes (W, 4, np, Np), or null
  float* S;             // Taylor partial sum (W, Np, np) complex64
  const double* win;    // walkers read for the FP64 update (W, N, n) complex128
  double* w;            // walkers written
  int W, N, n, Np, np_, nb;
  float c;              // Taylor coefficient sqrt(dt) / k
  int first, last;      // Taylor term position
  int tn_hi;            // the next kernel reads only the hi planes of tn (one-pass term): lo planes not stored
};

__device__ __forceinline__ bf16x8_t neg_bf16x8(bf16x8_t x) {
  u32x4_t u = __builtin_bit_cast(u32x4_t, x);
  u ^= (u32x4_t){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  return __builtin_bit_cast(bf16x8_t, u);
}

__device__ __forceinline__ unsigned pack2(bf16raw a, bf16raw b) { return (unsigned)a | ((unsigned)b << 16); }

// LDS offset (bf16 elements) of 16-byte segment seg of row r in a tile of BK-wide rows.  BK = 32: the
// swizzle of the split-BF16 GEMMs.  BK = 64 (128-byte rows): segment seg ^ ((r >> 1) & 7), so that the 16
// rows of a fragment read (one segment each) fall on 16 distinct 16-byte bank groups, and the two 16-lane
// halves of a 32-lane read (segments s, s + 1) on 32 distinct ones.
template <int BK>
__device__ __forceinline__ int prop_lds(int r, int seg) {
  if constexpr (BK == 32) {
    return b3_lds(r, seg);
  } else {
    return r * 64 + 8 * (seg ^ ((r >> 1) & 7));
  }
}

// REAL_A: product (E - I) . planes (one body); else V . planes (one Taylor term).
// VM (Taylor only): 0 V as complex64, split here; 1 V as split planes (hi, lo), three passes; 2 the hi
// planes of V and of the term only, one pass (bf16 products for the late, small terms).
// K advances BK = 64 per staged step with 8 waves (every thread stages a share of the A and B tiles) and
// 32 with 4 waves; either way each accumulator sees its MFMAs in ascending K steps of 32, so the result
// does not depend on BK.
template <bool REAL_A, int VM, int NW>
__device__ __forceinline__ void prop_b3_body(const PropB3P& p, const int b) {
  constexpr int BM = PB_BM, BN = PB_BN;
  constexpr int BK = NW == 8 ? 64 : 32;
  constexpr int NSEG = BK / 8;     // 16-byte segments per LDS row
  static_assert(64 * NW == BN * NSEG, "one B segment per thread per plane");
  constexpr int TI = NW == 8 ? 2 : 4;  // 16-row MFMA tiles per wave: waves of 64 x 32 (4 waves) or 32 x 32 (8)
  constexpr int PA = REAL_A ? 2 : 4;
  constexpr bool HIONLY = !REAL_A && VM == 2;  // lo planes neither loaded nor multiplied
  constexpr bool BHI = HIONLY;
  // planes held in LDS: the one-pass form keeps only re_hi, im_hi (slots 0, 1)
  constexpr int LA = HIONLY ? 2 : PA, LB = BHI ? 2 : 4;
  // 8 waves: 48 KB (one-pass) or 64 KB (one body); 4 waves: 48 KB.  Two workgroups per CU
  __shared__ __attribute__((aligned(16))) bf16raw As[LA][BM * BK];
  __shared__ __attribute__((aligned(16))) bf16raw Bs[LB][BN * BK];
  auto aslot = [](int q) { return HIONLY ? q >> 1 : q; };
  auto bslot = [](int q) { return BHI ? q >> 1 : q; };

  const int n_rt = p.Np / BM;
  const int tpw = n_rt * (p.np_ / BN);
  // b: the workgroup's index within this problem's grid
  const int grp = b >> 3;
  const int tile = grp % tpw;
  const long walker = (long)(grp / tpw) * 8 + (b & 7);
  if (walker >= p.W) return;
  const int rt = tile % n_rt, ct = tile / n_rt;
  const int m0 = rt * BM, n0 = ct * BN;
  const long tplane = (long)p.np_ * p.Np;
  const bf16raw* Bb = p.tp + walker * 4 * tplane + (long)n0 * p.Np;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int row0 = t / NSEG, sq = t % NSEG;  // B row (and E rows row0, row0 + 64), 16-byte segment

  // V staging: thread t reads stored rows 2 aq, 2 aq + 1 (128 contiguous bytes) of 8 x 8 block ablk
  const int ablk = t >> 2, aq = t & 3;
  const int rbl = ablk / NSEG, kbl = ablk % NSEG;
  const int rb = (m0 >> 3) + rbl;
  const float* Vb = REAL_A ? nullptr : p.vsym + walker * p.vcols * 2;
  const bf16raw* Eb = REAL_A ? p.ep + (long)m0 * p.Np : nullptr;
  bool up = true;

  u32x4_t est[REAL_A ? 4 : 1];
  f4_t ast[!REAL_A && VM == 0 ? 8 : 1];
  u32x4_t vhw[VM != 0 ? 4 : 1], vlw[VM == 1 ? 4 : 1];
  u32x4_t bst[4];
  auto load = [&](int k0) {
    if constexpr (REAL_A) {
      const long off0 = (long)row0 * p.Np + k0 + 8 * sq;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        est[2 * q] = *reinterpret_cast<const u32x4_t*>(Eb + q * (long)p.Np * p.Np + off0);
        est[2 * q + 1] = *reinterpret_cast<const u32x4_t*>(Eb + q * (long)p.Np * p.Np + off0 + 64L * p.Np);
      }
    } else {
      const int kb = (k0 >> 3) + kbl;
      up = rb <= kb;
      const bool inb = rb < p.nb && kb < p.nb;
      const long s = inb ? (up ? symblk_index(rb, kb, p.nb) : symblk_index(kb, rb, p.nb)) : 0;
      if constexpr (VM == 0) {
        const f4_t* src = reinterpret_cast<const f4_t*>(Vb + s * 128 + 32 * aq);
#pragma unroll
        for (int j = 0; j < 8; ++j) ast[j] = inb ? src[j] : f4_t{0.f, 0.f, 0.f, 0.f};
      } else {
        const long o = walker * p.vcols + s * 64 + 16 * aq;
        const u32x4_t* hsrc = reinterpret_cast<const u32x4_t*>(p.vh + o);
        const u32x4_t z4 = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int j = 0; j < 4; ++j) vhw[j] = inb ? hsrc[j] : z4;
        if constexpr (VM == 1) {
          const u32x4_t* lsrc = reinterpret_cast<const u32x4_t*>(p.vl + o);
#pragma unroll
          for (int j = 0; j < 4; ++j) vlw[j] = inb ? lsrc[j] : z4;
        }
      }
    }
    const long offb = (long)row0 * p.Np + k0 + 8 * sq;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (BHI && (q & 1)) continue;
      bst[q] = *reinterpret_cast<const u32x4_t*>(Bb + q * tplane + offb);
    }
  };
  auto store = [&]() {
    if constexpr (REAL_A) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        *reinterpret_cast<u32x4_t*>(&As[q][prop_lds<BK>(row0, sq)]) = est[2 * q];
        *reinterpret_cast<u32x4_t*>(&As[q][prop_lds<BK>(row0 + 64, sq)]) = est[2 * q + 1];
      }
    } else {
      // complex c = 8 srl + sc of the two stored rows: (re, im) = ast[c / 2][2 (c % 2)], [2 (c % 2) + 1]
      bf16raw h[4][16];
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        if constexpr (VM == 0) {
          split_bf16_raw(ast[c >> 1][2 * (c & 1)], h[0][c], h[1][c]);
          split_bf16_raw(ast[c >> 1][2 * (c & 1) + 1], h[2][c], h[3][c]);
        } else {
          const unsigned wh = vhw[c >> 2][c & 3];
          h[0][c] = (bf16raw)(wh & 0xffffu);
          h[2][c] = (bf16raw)(wh >> 16);
          if constexpr (VM == 1) {
            const unsigned wl = vlw[c >> 2][c & 3];
            h[1][c] = (bf16raw)(wl & 0xffffu);
            h[3][c] = (bf16raw)(wl >> 16);
          } else {
            h[1][c] = h[3][c] = 0;
          }
        }
      }
      if (up) {  // stored row 2 aq + srl is tile row 8 rbl + 2 aq + srl, k = 8 kbl + sc
#pragma unroll
        for (int srl = 0; srl < 2; ++srl)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (HIONLY && (q & 1)) continue;
            const bf16raw* v = &h[q][8 * srl];
            *reinterpret_cast<u32x4_t*>(&As[aslot(q)][prop_lds<BK>(8 * rbl + 2 * aq + srl, kbl)]) =
                (u32x4_t){pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7])};
          }
      } else {  // mirrored block: stored (2 aq + srl, sc) is tile row 8 rbl + sc, k = 8 kbl + 2 aq + srl
#pragma unroll
        for (int sc = 0; sc < 8; ++sc)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (HIONLY && (q & 1)) continue;
            *reinterpret_cast<unsigned*>(&As[aslot(q)][prop_lds<BK>(8 * rbl + sc, kbl) + 2 * aq]) =
                pack2(h[q][sc], h[q][8 + sc]);
          }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (BHI && (q & 1)) continue;
      *reinterpret_cast<u32x4_t*>(&Bs[bslot(q)][prop_lds<BK>(row0, sq)]) = bst[q];
    }
  };

  f4_t acc[2][TI][2];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[q][i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  const int nk = p.Np / BK;
  load(0);
  for (int kc = 0; kc < nk; ++kc) {
    __syncthreads();
    store();
    __syncthreads();
    if (kc + 1 < nk) load((kc + 1) * BK);
#pragma unroll
    for (int u = 0; u < BK / 32; ++u) {
      const int seg = (lane >> 4) + 4 * u;
      // B fragments [j][plane], plane 0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo
      bf16x8_t bf[2][4];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = wn * 32 + j * 16 + (lane & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (BHI && (q & 1)) continue;
          bf[j][q] = *reinterpret_cast<const bf16x8_t*>(&Bs[bslot(q)][prop_lds<BK>(col, seg)]);
        }
      }
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        const int row = wm * 16 * TI + i * 16 + (lane & 15);
        bf16x8_t a[PA];
#pragma unroll
        for (int q = 0; q < PA; ++q) {
          if (HIONLY && (q & 1)) continue;
          a[q] = *reinterpret_cast<const bf16x8_t*>(&As[aslot(q)][prop_lds<BK>(row, seg)]);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          f4_t cr = acc[0][i][j], ci = acc[1][i][j];
          if constexpr (REAL_A) {  // re += E.Br, im += E.Bi
            cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bf[j][0], cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][1], cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][0], cr, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bf[j][2], ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][3], ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][2], ci, 0, 0, 0);
          } else {
            const bf16x8_t nh = neg_bf16x8(bf[j][2]);
            // re += Ar.Br - Ai.Bi
            if constexpr (!HIONLY) cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bf[j][0], cr, 0, 0, 0);
            if constexpr (!BHI) cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][1], cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][0], cr, 0, 0, 0);
            if constexpr (!HIONLY) cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[3], nh, cr, 0, 0, 0);
            if constexpr (!BHI) cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], neg_bf16x8(bf[j][3]), cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], nh, cr, 0, 0, 0);
            // im += Ar.Bi + Ai.Br
            if constexpr (!HIONLY) ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bf[j][2], ci, 0, 0, 0);
            if constexpr (!BHI) ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][3], ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bf[j][2], ci, 0, 0, 0);
            if constexpr (!HIONLY) ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[3], bf[j][0], ci, 0, 0, 0);
            if constexpr (!BHI) ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], bf[j][1], ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], bf[j][0], ci, 0, 0, 0);
          }
          acc[0][i][j] = cr;
          acc[1][i][j] = ci;
        }
      }
    }
  }

  // epilogue.  Lane holds rows 4 (l / 16) + r, column l % 16 of each 16 x 16 tile.
  // The loads of a group of EI row tiles (old running sums, FP64 walkers) are issued together, before any
  // of the group's stores: the counter a wave waits on for a load also counts its earlier stores, so loads
  // placed between stores (as a per-element read-modify-write compiles) would each wait for every store
  // before them.  EI keeps the loaded values within the registers the main loop leaves free.
  // Each array is addressed from one per-lane base (row gib, column gpb: r = 0 of row tile 0, column tile 0)
  // plus offsets that are the same for the whole wave.
  const bool to_w = REAL_A || p.last;
  const int gib = m0 + wm * 16 * TI + 4 * (lane >> 4), gpb = n0 + wn * 32 + (lane & 15);
  const long sld = p.np_, wld = p.n;
  float2* Sb = REAL_A ? nullptr : reinterpret_cast<float2*>(p.S) + (walker * p.Np + gib) * sld + gpb;
  const double2* Wi = to_w ? reinterpret_cast<const double2*>(p.win) + (walker * p.N + gib) * wld + gpb : nullptr;
  double2* Wo = to_w ? reinterpret_cast<double2*>(p.w) + (walker * p.N + gib) * wld + gpb : nullptr;
  bf16raw* Tb = p.tn ? p.tn + walker * 4 * tplane + (long)gpb * p.Np + gib : nullptr;
  constexpr int EI = REAL_A ? TI : (NW == 8 ? 1 : 2);
  static_assert(TI % EI == 0, "row tiles per epilogue group");
#pragma unroll
  for (int i0 = 0; i0 < TI; i0 += EI) {
    float2 so[EI][2][4];
    double2 wo[EI][2][4];
#pragma unroll
    for (int ii = 0; ii < EI; ++ii)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int di = 16 * (i0 + ii) + r;
          so[ii][j][r] = make_float2(0.f, 0.f);
          if (!REAL_A && !p.first) so[ii][j][r] = Sb[di * sld + 16 * j];
          wo[ii][j][r] = make_double2(0.0, 0.0);
          if (to_w && gib + di < p.N && gpb + 16 * j < p.n) wo[ii][j][r] = Wi[di * wld + 16 * j];
        }
#pragma unroll
    for (int ii = 0; ii < EI; ++ii)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int i = i0 + ii;
        bf16raw hr[4], lr[4], hi[4], li[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int di = 16 * i + r;
          float tre, tim;
          if constexpr (REAL_A) {
            tre = acc[0][i][j][r];
            tim = acc[1][i][j][r];
          } else {  // T = i c acc
            tre = -p.c * acc[1][i][j][r];
            tim = p.c * acc[0][i][j][r];
          }
          float pr = tre, pi = tim;  // value whose planes are written
          if (!REAL_A) {
            const float2 o = so[ii][j][r];
            if (!p.last) {
              Sb[di * sld + 16 * j] = make_float2(o.x + tre, o.y + tim);
            } else {
              tre += o.x;
              tim += o.y;
            }
          }
          if (to_w) {
            pr = pi = 0.f;
            if (gib + di < p.N && gpb + 16 * j < p.n) {
              const double2 o = wo[ii][j][r];
              const double2 v = make_double2(o.x + (double)tre, o.y + (double)tim);
              Wo[di * wld + 16 * j] = v;
              pr = (float)v.x;
              pi = (float)v.y;
            }
          }
          split_bf16_raw(pr, hr[r], lr[r]);
          split_bf16_raw(pi, hi[r], li[r]);
        }
        if (Tb) {  // 4 consecutive rows of row gp of the transposed planes: one 8-byte store per plane
          bf16raw* o = Tb + (long)(16 * j) * p.Np + 16 * i;
          typedef unsigned short us4 __attribute__((ext_vector_type(4)));
          *reinterpret_cast<us4*>(o + 0 * tplane) = (us4){hr[0], hr[1], hr[2], hr[3]};
          *reinterpret_cast<us4*>(o + 2 * tplane) = (us4){hi[0], hi[1], hi[2], hi[3]};
          if (!p.tn_hi) {
            *reinterpret_cast<us4*>(o + 1 * tplane) = (us4){lr[0], lr[1], lr[2], lr[3]};
            *reinterpret_cast<us4*>(o + 3 * tplane) = (us4){li[0], li[1], li[2], li[3]};
          }
        }
      }
  }
}

template <bool REAL_A, int VM = 0>
__global__ __launch_bounds__((64 * PropWaves<REAL_A, VM>::value), (PropWaves<REAL_A, VM>::value / 2)) void prop_b3_kernel(PropB3P p) {
  prop_b3_body<REAL_A, VM, PropWaves<REAL_A, VM>::value>(p, blockIdx.x);
}

// The propagation kernels of many walker sets in one launch: problem i owns workgroups start[i] ..
// start[i+1] - 1 (a multiple of 8 each, so a workgroup's XCD is that of its place in its own grid) and
// runs exactly the workgroup body of its own launch.
__device__ __forceinline__ int b3_problem_of(const int* __restrict__ start, int nprob, int g) {
  int lo = 0, hi = nprob;  // start[lo] <= g < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

template <bool REAL_A, int VM>
__global__ __launch_bounds__((64 * PropWaves<REAL_A, VM>::value), (PropWaves<REAL_A, VM>::value / 2)) void prop_b3_grouped_kernel(
    const PropB3P* __restrict__ probs, const int* __restrict__ start, int nprob) {
  const int pi = b3_problem_of(start, nprob, blockIdx.x);
  const PropB3P p = probs[pi];
  prop_b3_body<REAL_A, VM, PropWaves<REAL_A, VM>::value>(p, blockIdx.x - start[pi]);
}

// Transposed planes (W, 4, np, Np) of walkers w (W, N, n) complex128, zero padded.  Grid (Np/64, np/64, W),
// 256 threads; a 64 x 64 tile goes through LDS (as float2: the split starts from the FP32 value, and 33 KB
// leave room for four workgroups per CU) so that reads run along p, and each thread then converts 8
// consecutive orbitals of one column and writes them as one 16-byte store per plane.
__device__ __forceinline__ void split_transpose_w_body(const double* __restrict__ w, bf16raw* __restrict__ tp, int N,
                                                       int n, int Np, int np_, int bx, int by, long bz) {
  __shared__ float2 tile[64][65];
  const int i0 = bx * 64, p0 = by * 64;
  const long b = bz;
  const int t = threadIdx.x;
  const double2* wb = reinterpret_cast<const double2*>(w) + b * (long)N * n;
  for (int e = t; e < 64 * 64; e += 256) {
    const int ii = e >> 6, pp = e & 63;  // read along p
    const int gi = i0 + ii, gp = p0 + pp;
    const double2 v = (gi < N && gp < n) ? wb[(long)gi * n + gp] : make_double2(0.0, 0.0);
    tile[ii][pp] = make_float2((float)v.x, (float)v.y);
  }
  __syncthreads();
  const long plane = (long)np_ * Np;
  bf16raw* base = tp + b * 4 * plane;
  typedef unsigned short us8 __attribute__((ext_vector_type(8)));
  for (int e = t; e < 64 * 8; e += 256) {
    const int pp = e >> 3, i8 = (e & 7) * 8;  // column pp, orbitals i8 .. i8 + 7 of the tile
    us8 rh, rl, ih, il;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float2 v = tile[i8 + q][pp];
      bf16raw h, l;
      split_bf16_raw(v.x, h, l);
      rh[q] = h;
      rl[q] = l;
      split_bf16_raw(v.y, h, l);
      ih[q] = h;
      il[q] = l;
    }
    const long o = (long)(p0 + pp) * Np + i0 + i8;
    *reinterpret_cast<us8*>(base + o) = rh;
    *reinterpret_cast<us8*>(base + plane + o) = rl;
    *reinterpret_cast<us8*>(base + 2 * plane + o) = ih;
    *reinterpret_cast<us8*>(base + 3 * plane + o) = il;
  }
}

__global__ __launch_bounds__(256) void split_transpose_w_kernel(const double* __restrict__ w, bf16raw* __restrict__ tp,
                                                                int N, int n, int Np, int np_) {
  split_transpose_w_body(w, tp, N, n, Np, np_, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Grouped: problem i (a PropB3P with win = walkers, tn = planes, W, N, n, Np, np_) owns workgroups
// start[i] .. start[i+1] - 1, laid out as its own grid (x fastest, then y, then z = walker).
__global__ __launch_bounds__(256) void split_transpose_grouped_kernel(const PropB3P* __restrict__ probs,
                                                                      const int* __restrict__ start, int nprob) {
  const int pi = b3_problem_of(start, nprob, blockIdx.x);
  const PropB3P p = probs[pi];
  const int local = blockIdx.x - start[pi];
  const int nx = p.Np / 64, ny = p.np_ / 64;
  split_transpose_w_body(p.win, p.tn, p.N, p.n, p.Np, p.np_, local % nx, (local / nx) % ny, local / (nx * ny));
}

// (E - I) split into planes [hi, lo] of (Np, Np), zero padded.
__global__ void onebody_b3_planes_kernel(const double* __restrict__ e, int N, int Np, bf16raw* __restrict__ out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)Np * Np;
  if (idx >= total) return;
  const int i = (int)(idx / Np), k = (int)(idx - (idx / Np) * Np);
  const double v = (i < N && k < N) ? e[(long)i * N + k] - (i == k ? 1.0 : 0.0) : 0.0;
  bf16raw h, l;
  split_bf16_raw((float)v, h, l);
  out[idx] = h;
  out[total + idx] = l;
}

inline int pad_up(int x, int m) { return (x + m - 1) / m * m; }

struct B3Buffers {
  bf16raw* t0;
  bf16raw* t1;
  float* S;
};

size_t b3_workspace(int W, int N, int n) {
  const size_t Np = pad_up(N, PB_BM), np_ = pad_up(n, PB_BN);
  return 2 * align256((size_t)W * 4 * np_ * Np * 2) + align256((size_t)W * Np * np_ * 8);
}

template <bool REAL_A, int VM = 0>
int launch_prop(const PropB3P& P, hipStream_t s) {
  const int tpw = (P.Np / PB_BM) * (P.np_ / PB_BN);
  const long nwg = (long)((P.W + 7) / 8) * 8 * tpw;
  if (nwg > 0x7fffffffL) return fail(TROT_ERR_INVALID, "propagation b3: too many walkers");
  constexpr int threads = 64 * PropWaves<REAL_A, VM>::value;
  hipLaunchKernelGGL((prop_b3_kernel<REAL_A, VM>), dim3((unsigned)nwg), dim3(threads), 0, s, P);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

// V of one call: complex64 (c64), or split planes (vh, vl) with terms k >= hi_from on the hi plane only.
struct VSrc {
  const float* c64;
  const unsigned* vh;
  const unsigned* vl;
  int hi_from;  // terms k >= hi_from: one bf16 pass on the hi planes of V and of the term
};

// One launch of the propagation: its kind and parameters.  SPLIT: walkers win -> planes tn; ONEBODY: the
// (E - I) product; TERM_C64 / TERM3 / TERM1: a Taylor term with V complex64, split V three passes, split V
// one pass; COPY: win -> w (no one-body factors and no terms).
enum { B3_SPLIT = 0, B3_ONEBODY = 1, B3_TERM_C64 = 2, B3_TERM3 = 3, B3_TERM1 = 4, B3_COPY = 5 };
struct B3Stage {
  int kind;
  PropB3P P;
};

long b3_stage_workgroups(const B3Stage& st) {
  const PropB3P& P = st.P;
  if (st.kind == B3_SPLIT) return (long)(P.Np / 64) * (P.np_ / 64) * P.W;
  if (st.kind == B3_COPY) return 0;
  return (long)((P.W + 7) / 8) * 8 * (P.Np / PB_BM) * (P.np_ / PB_BN);
}

// w_out <- [E .] T_K(i sqrt(dt) V) [. E] w_in; one-body factors when ep is non-null, as a list of launches.
//
// The split of w_in into planes and both one-body factors run once over all walkers (E is shared, so
// these launches are large and short-lived per walker); the Taylor series runs over walker chunks whose
// V, term planes and partial sums (about 1.8 MB per walker at N = 512, n = 64) stay within the 256 MB
// Infinity Cache across the whole series, while each launch still has at least 512 workgroups (two per
// CU).  Measured at N = 512, n = 64, W = 2048 with the one-body factors inside the chunks: 128 walkers per
// chunk 10.5 ms, all at once 11.9 ms, 64 per chunk 14.3 ms (fewer workgroups than CUs can hold).
int b3_stages(const VSrc& v, const bf16raw* ep, const double* w_in, double* w_out, int W, int N, int n,
              double sqrt_dt, int n_terms, void* ws, size_t ws_bytes, std::vector<B3Stage>& out) {
  const int Np = pad_up(N, PB_BM), np_ = pad_up(n, PB_BN), nb = (N + 7) / 8;
  Carve cv{(char*)ws, ws_bytes, 0};
  bf16raw* bufs[2] = {cv.take<bf16raw>((size_t)W * 4 * np_ * Np), cv.take<bf16raw>((size_t)W * 4 * np_ * Np)};
  float* S = cv.take<float>((size_t)W * Np * np_ * 2);
  if (!bufs[0] || !bufs[1] || !S) return fail(TROT_ERR_WORKSPACE, "propagation b3: workspace");
  const long vcols = symblk_count(nb) * 64;
  PropB3P base{};
  base.vcols = vcols;
  base.ep = ep;
  base.W = W;
  base.N = N;
  base.n = n;
  base.Np = Np;
  base.np_ = np_;
  base.nb = nb;
  if (!ep && n_terms == 0) {
    if (w_out != w_in) {
      B3Stage st{B3_COPY, base};
      st.P.win = w_in;
      st.P.w = w_out;
      out.push_back(st);
    }
    return TROT_OK;
  }
  {
    B3Stage st{B3_SPLIT, base};
    st.P.win = w_in;
    st.P.tn = bufs[0];
    out.push_back(st);
  }
  int cur = 0;
  if (ep) {  // first one-body factor, all walkers: w_out = w_in + (E - I) w_in, with its planes
    B3Stage st{B3_ONEBODY, base};
    st.P.tp = bufs[0];
    st.P.tn = bufs[1];
    st.P.win = w_in;
    st.P.w = w_out;
    out.push_back(st);
    cur = 1;
  } else if (w_out != w_in) {
    B3Stage st{B3_COPY, base};
    st.P.win = w_in;
    st.P.w = w_out;
    out.push_back(st);
  }
  const double per_walker = 8.0 * vcols + 2.0 * 4 * np_ * Np * 2 + 8.0 * Np * np_;
  const int tpw = (Np / PB_BM) * (np_ / PB_BN);
  int chunk = (int)(200e6 / per_walker);
  chunk = chunk < (512 + tpw - 1) / tpw ? (512 + tpw - 1) / tpw : chunk;
  chunk = (chunk + 7) / 8 * 8;
  if (chunk > W) chunk = W;
  const size_t plane_w = (size_t)4 * np_ * Np;  // bf16 words of one walker's planes
  int cur_end = cur;
  for (int c0 = 0; c0 < W; c0 += chunk) {
    const int wc = W - c0 < chunk ? W - c0 : chunk;
    PropB3P P = base;
    P.W = wc;
    if (v.c64) P.vsym = v.c64 + (size_t)c0 * vcols * 2;
    if (v.vh) P.vh = v.vh + (size_t)c0 * vcols;
    if (v.vl) P.vl = v.vl + (size_t)c0 * vcols;
    P.S = S + (size_t)c0 * Np * np_ * 2;
    P.w = w_out + (size_t)c0 * N * n * 2;
    P.win = P.w;
    int c = cur;
    for (int k = 1; k <= n_terms; ++k) {
      P.tp = bufs[c] + c0 * plane_w;
      P.first = k == 1;
      P.last = k == n_terms;
      P.tn = (P.last && !ep) ? nullptr : bufs[c ^ 1] + c0 * plane_w;
      // the next term, when it is one bf16 pass on the hi planes, never reads the lo planes of this output
      P.tn_hi = !v.c64 && k < n_terms && k + 1 >= v.hi_from;
      P.c = (float)(sqrt_dt / k);
      out.push_back(B3Stage{v.c64 ? B3_TERM_C64 : (k >= v.hi_from ? B3_TERM1 : B3_TERM3), P});
      c ^= 1;
    }
    cur_end = c;
  }
  if (ep) {  // second one-body factor, all walkers, on the planes of the propagated walkers
    B3Stage st{B3_ONEBODY, base};
    st.P.tp = bufs[cur_end];
    st.P.tn = nullptr;
    st.P.win = w_out;
    st.P.w = w_out;
    out.push_back(st);
  }
  return TROT_OK;
}

int launch_stage(const B3Stage& st, hipStream_t s) {
  const PropB3P& P = st.P;
  switch (st.kind) {
    case B3_SPLIT:
      hipLaunchKernelGGL(split_transpose_w_kernel, dim3(P.Np / 64, P.np_ / 64, P.W), dim3(256), 0, s, P.win, P.tn, P.N,
                         P.n, P.Np, P.np_);
      HIPCHK(hipGetLastError());
      return TROT_OK;
    case B3_ONEBODY: return launch_prop<true>(P, s);
    case B3_TERM_C64: return launch_prop<false, 0>(P, s);
    case B3_TERM3: return launch_prop<false, 1>(P, s);
    case B3_TERM1: return launch_prop<false, 2>(P, s);
    default:
      HIPCHK(hipMemcpyAsync(P.w, P.win, (size_t)P.W * P.N * P.n * 16, hipMemcpyDeviceToDevice, s));
      return TROT_OK;
  }
}

int run_b3(const VSrc& v, const bf16raw* ep, const double* w_in, double* w_out, int W, int N, int n,
           double sqrt_dt, int n_terms, void* ws, size_t ws_bytes, hipStream_t s) {
  std::vector<B3Stage> stages;
  if (int rc = b3_stages(v, ep, w_in, w_out, W, N, n, sqrt_dt, n_terms, ws, ws_bytes, stages)) return rc;
  for (const B3Stage& st : stages)
    if (int rc = launch_stage(st, s)) return rc;
  return TROT_OK;
}

int check_sizes(const char* who, int W, int N, int n, int n_terms) {
  if (W < 1 || W > 65535 || N < 1 || n < 1 || n_terms < 0) {
    char msg[128];
    snprintf(msg, sizeof(msg), "%s: bad sizes", who);
    return fail(TROT_ERR_INVALID, msg);
  }
  return TROT_OK;
}

}  // namespace

extern "C" {

size_t trot_taylor_b3_workspace(int W, int N, int n) { return b3_workspace(W, N, n); }

// w (complex128, in/out) <- sum_{k=0..K} (i sqrt(dt) V)^k / k! w with V given in sym-block storage as
// complex64 (W, trot_symblk_cols(N)).  Products on split-BF16 MFMA; terms summed in FP32, added to w in FP64.
int trot_taylor_apply_b3_symblk(const float* vhs_symblk, double* w, int W, int N, int n, double sqrt_dt, int n_terms,
                                void* ws, size_t ws_bytes, void* stream) {
  if (!vhs_symblk || !w) return fail(TROT_ERR_INVALID, "taylor_apply_b3: null pointer");
  if (int rc = check_sizes("taylor_apply_b3", W, N, n, n_terms)) return rc;
  if (ws_bytes < b3_workspace(W, N, n)) return fail(TROT_ERR_WORKSPACE, "taylor_apply_b3: workspace");
  return run_b3(VSrc{vhs_symblk, nullptr, nullptr, 0}, nullptr, w, w, W, N, n, sqrt_dt, n_terms, ws, ws_bytes,
                (hipStream_t)stream);
}

size_t trot_onebody_b3_planes_bytes(int N) {
  const size_t Np = pad_up(N, PB_BM);
  return 2 * Np * Np * 2;
}

int trot_onebody_b3_planes(const double* expm, int N, void* planes, void* stream) {
  if (!expm || !planes || N < 1) return fail(TROT_ERR_INVALID, "onebody_b3_planes: bad arguments");
  const int Np = pad_up(N, PB_BM);
  const long total = (long)Np * Np;
  hipLaunchKernelGGL(onebody_b3_planes_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, expm, N, Np, (bf16raw*)planes);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

int trot_trotter_b3_symblk(const float* vhs_symblk, const void* onebody_planes, const double* w_in, double* w_out,
                           int W, int N, int n, double sqrt_dt, int n_terms, void* ws, size_t ws_bytes, void* stream) {
  if (!vhs_symblk || !onebody_planes || !w_in || !w_out) return fail(TROT_ERR_INVALID, "trotter_b3: null pointer");
  if (int rc = check_sizes("trotter_b3", W, N, n, n_terms)) return rc;
  if (ws_bytes < b3_workspace(W, N, n)) return fail(TROT_ERR_WORKSPACE, "trotter_b3: workspace");
  return run_b3(VSrc{vhs_symblk, nullptr, nullptr, 0}, (const bf16raw*)onebody_planes, w_in, w_out, W, N, n,
                sqrt_dt, n_terms, ws, ws_bytes, (hipStream_t)stream);
}

// As trot_trotter_b3_symblk with V as split planes [hi | lo], each (W, trot_symblk_cols(N)) words of
// bf16 re | bf16 im << 16 (trot_gemm_b3 output kind 2); terms k >= hi_from are one bf16 pass on the hi
// planes of V and of the term.
int trot_trotter_b3v(const void* v_planes, int hi_from, const void* onebody_planes, const double* w_in,
                     double* w_out, int W, int N, int n, double sqrt_dt, int n_terms, void* ws, size_t ws_bytes,
                     void* stream) {
  if (!v_planes || !onebody_planes || !w_in || !w_out) return fail(TROT_ERR_INVALID, "trotter_b3v: null pointer");
  if (int rc = check_sizes("trotter_b3v", W, N, n, n_terms)) return rc;
  if (hi_from < 1) return fail(TROT_ERR_INVALID, "trotter_b3v: hi_from must be >= 1");
  if (ws_bytes < b3_workspace(W, N, n)) return fail(TROT_ERR_WORKSPACE, "trotter_b3v: workspace");
  const long vcols = symblk_count((N + 7) / 8) * 64;
  const unsigned* vh = (const unsigned*)v_planes;
  return run_b3(VSrc{nullptr, vh, vh + (size_t)W * vcols, hi_from}, (const bf16raw*)onebody_planes, w_in, w_out, W,
                N, n, sqrt_dt, n_terms, ws, ws_bytes, (hipStream_t)stream);
}


size_t trot_prop_b3_stage_bytes(void) { return sizeof(PropB3P); }

// The launches trot_trotter_b3v makes, as descriptors for grouped launches over many walker sets: kinds[i]
// (TROT_B3_STAGE_*), workgroups[i] and descs (max_stages * trot_prop_b3_stage_bytes() bytes) of each, in
// order.  Returns the number of launches, or an error.  Nothing is launched; ws must stay allocated while
// the descriptors are used.
long trot_trotter_b3v_stages(const void* v_planes, int hi_from, const void* onebody_planes, const double* w_in,
                             double* w_out, int W, int N, int n, double sqrt_dt, int n_terms, void* ws,
                             size_t ws_bytes, int max_stages, int* kinds, long* workgroups, void* descs) {
  if (!v_planes || !onebody_planes || !w_in || !w_out || !kinds || !workgroups || !descs)
    return fail(TROT_ERR_INVALID, "trotter_b3v_stages: null pointer");
  if (int rc = check_sizes("trotter_b3v_stages", W, N, n, n_terms)) return rc;
  if (hi_from < 1) return fail(TROT_ERR_INVALID, "trotter_b3v_stages: hi_from must be >= 1");
  if (ws_bytes < b3_workspace(W, N, n)) return fail(TROT_ERR_WORKSPACE, "trotter_b3v_stages: workspace");
  const long vcols = symblk_count((N + 7) / 8) * 64;
  const unsigned* vh = (const unsigned*)v_planes;
  std::vector<B3Stage> stages;
  if (int rc = b3_stages(VSrc{nullptr, vh, vh + (size_t)W * vcols, hi_from}, (const bf16raw*)onebody_planes, w_in,
                         w_out, W, N, n, sqrt_dt, n_terms, ws, ws_bytes, stages))
    return rc;
  if ((long)stages.size() > max_stages) return fail(TROT_ERR_INVALID, "trotter_b3v_stages: more launches than max_stages");
  for (size_t i = 0; i < stages.size(); ++i) {
    kinds[i] = stages[i].kind;
    workgroups[i] = b3_stage_workgroups(stages[i]);
    reinterpret_cast<PropB3P*>(descs)[i] = stages[i].P;
  }
  return (long)stages.size();
}

// One launch of a propagation stage kind over many walker sets: probs (device) holds nprob descriptors from
// trot_trotter_b3v_stages of that kind, start (device) the nprob + 1 prefix sums of their workgroup counts.
// Each workgroup runs exactly the body it runs in its own set's launch (bit-identical results).
int trot_prop_b3_grouped(int kind, const void* probs, const int* start, int nprob, int total_wg, void* stream) {
  if (!probs || !start || nprob < 1 || total_wg < 1) return fail(TROT_ERR_INVALID, "prop_b3_grouped: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const PropB3P* P = (const PropB3P*)probs;
  constexpr int nw_ob = 64 * PropWaves<true, 0>::value, nw_c64 = 64 * PropWaves<false, 0>::value;
  constexpr int nw_t3 = 64 * PropWaves<false, 1>::value, nw_t1 = 64 * PropWaves<false, 2>::value;
  switch (kind) {
    case B3_SPLIT:
      hipLaunchKernelGGL(split_transpose_grouped_kernel, dim3(total_wg), dim3(256), 0, s, P, start, nprob);
      break;
    case B3_ONEBODY:
      hipLaunchKernelGGL((prop_b3_grouped_kernel<true, 0>), dim3(total_wg), dim3(nw_ob), 0, s, P, start, nprob);
      break;
    case B3_TERM_C64:
      hipLaunchKernelGGL((prop_b3_grouped_kernel<false, 0>), dim3(total_wg), dim3(nw_c64), 0, s, P, start, nprob);
      break;
    case B3_TERM3:
      hipLaunchKernelGGL((prop_b3_grouped_kernel<false, 1>), dim3(total_wg), dim3(nw_t3), 0, s, P, start, nprob);
      break;
    case B3_TERM1:
      hipLaunchKernelGGL((prop_b3_grouped_kernel<false, 2>), dim3(total_wg), dim3(nw_t1), 0, s, P, start, nprob);
      break;
    default: return fail(TROT_ERR_UNSUPPORTED, "prop_b3_grouped: this stage kind has no grouped launch");
  }
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

}  // extern "C"
