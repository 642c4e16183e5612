// Host-side launchers of the trot_amd kernels (definitions next to the kernels).
#pragma once

#include <hip/hip_runtime.h>

#define TROT_MAX_NOCC 64       /* register-resident inverse, Cholesky-QR, exchange blocks */
#define TROT_MAX_NOCC_BIG 256  /* the general inverse, tiled exchange and Gram-Schmidt QR above 64 */

hipError_t launch_inv_det(const double* m, double* minv, double* det, int n, int W, hipStream_t s);

// grouped launches over many walker sets (layouts shared with the C ABI, include/trot_amd.h)
struct InvProb {
  const double* m;
  double* minv;
  double* det;
  long n;
};
struct CopyProb {
  const void* src;
  void* dst;
  long bytes;
};
hipError_t launch_inv_det_grouped(const InvProb* probs, const int* start, int nprob, int total_w, hipStream_t s);
hipError_t launch_copy_grouped(const CopyProb* probs, int nprob, long max_bytes, hipStream_t s);
hipError_t launch_overlap_combine(const double* det_a, const double* det_b, double* ovlp, int W, hipStream_t s);
// grouped launches over many walker sets (descriptor tables in device memory)
struct SumProb {  // dst = (accumulate ? dst : 0) + scale * sum_b src[b]  (len doubles, nb slices)
  const double* src;
  double* dst;
  long len;
  int nb;
  int accumulate;
  double scale;
};
struct ShiftProb {
  const double* fields;
  const double* fb;
  const double* mf;
  double* xbar;
  double* shift;
  double* fbt;
  double sqrt_dt;
  int G;
};
struct UpdProb {
  const double* det_a;
  const double* det_b;  // null: restricted walkers (overlap det_a^2)
  double* ovlp;         // old overlaps in, new overlaps out
  const double* shift;
  const double* fbt;
  double* weights;
  double* scal;
  long long* nodes;
  int W;
  double dt, sqrt_dt, h0p_re, h0p_im, w_floor, w_cap, damping;
};
hipError_t launch_shift_grouped(const ShiftProb* probs, const int* start, int nprob, int total_w, hipStream_t s);
hipError_t launch_sum_grouped(const SumProb* probs, const int* start, int nprob, int total_blocks, hipStream_t s);
hipError_t launch_update_grouped(const UpdProb* probs, int nprob, hipStream_t s);
hipError_t launch_shift(const double* fields, const double* fb, const double* mf, double sqrt_dt, int W, int G,
                        double* xbar, float* xbar32, double* shift, double* fbt, hipStream_t s);
hipError_t launch_weight_update(double* weights, const double* ov_old, const double* ov_new, const double* shift,
                                const double* fbt, double* scal, long long* nodes, int W, double dt, double sqrt_dt,
                                double h0p_re, double h0p_im, double w_floor, double w_cap, double damping,
                                hipStream_t s);
hipError_t launch_chol_rinv(const double* gram, double* rinv, int n, int W, hipStream_t s, double* dprod = nullptr,
                            int accumulate = 0, int* flags = nullptr);
hipError_t launch_qr_fallback(const double* src, double* q, const int* flags, int N, int n, int W, double* rinv,
                              double* dprod, hipStream_t s);
// zeta_dev: comb offset read on the device when non-null (else zeta)
hipError_t launch_sr_index(const double* weights, double* cw, int* idx, double* w_new, int W, double zeta,
                           const double* zeta_dev, hipStream_t s);
hipError_t launch_fill_pattern(double* x, long n, hipStream_t s);
hipError_t launch_gather_rows(const double* src, double* dst, const int* idx, int W, long row, hipStream_t s);
hipError_t launch_energy(const double* rc_a, const double* rc_b, const double* gt_a, const double* gt_b,
                         const double* h1t_a, const double* h1t_b, int na, int nb, int N, int G, int W, double h0,
                         int gb, double* partial, double* energy, hipStream_t s);
hipError_t launch_energy_lp(const float* rc_a, const float* rc_b, const double* gt_a, const double* gt_b,
                            const double* h1t_a, const double* h1t_b, int na, int nb, int N, int G, int W, double h0,
                            int gb, double* partial, double* energy, int mode, hipStream_t s);
hipError_t launch_energy_finalize(const double* gt_a, const double* gt_b, const double* h1t_a, const double* h1t_b,
                                  int na, int nb, int N, double h0, const double* partial, int nblk, double* energy,
                                  int W, hipStream_t s);
hipError_t launch_mf_trace(const double* chol, const double* dm, int G, int N, double* mf, hipStream_t s);
hipError_t launch_sum_batches(const double* src, double* dst, int nb, long len, double scale, int accumulate,
                              hipStream_t s);
hipError_t launch_block_energy(const double* energy, double* weights, double* scal, double* out, int W,
                               double thresh, double alpha, hipStream_t s);
hipError_t launch_pack_symblk(const double* chol, void* out, int out_f32, int G, int N, hipStream_t s);
hipError_t launch_c128_to_c64(const double* src, float* dst, long n, hipStream_t s);
