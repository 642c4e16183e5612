"""ctypes binding of libtrot_amd.so (include/trot_amd.h) with host-side shape checks.

The library is the only compute path: if it is missing, importing the
compute entry points raises.  There is no CPU or PyTorch fallback.  Every
wrapper checks device, dtype, contiguity and shapes of its tensors against
what the kernel and its launch grid assume before calling into HIP.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libtrot_amd.so"

TROT_OK = 0
MAX_NOCC = 64  # the register-resident inverse, Cholesky-QR, split-BF16 exchange blocks, observables, CISD
MAX_NOCC_BIG = 256  # the general inverse, tiled FP64 / FP32 exchange and Gram-Schmidt QR above 64


class TrotError(RuntimeError):
    """A trot_amd C entry point returned a non-zero status."""


_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_l = ctypes.c_long
_d = ctypes.c_double
_sz = ctypes.c_size_t
_u64 = ctypes.c_uint64

_SIGS = {
    "trot_version": (ctypes.c_char_p, []),
    "trot_last_error": (ctypes.c_char_p, []),
    "trot_overlap_green_workspace": (_sz, [_i, _i, _i]),
    "trot_overlap_green": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _sz, _vp]),
    "trot_overlap_combine": (_i, [_vp, _vp, _vp, _i, _vp]),
    "trot_force_bias": (_i, [_vp, _vp, _i, _i, _l, _d, _i, _vp, _vp]),
    "trot_force_bias_workspace": (_sz, [_i, _i, _l]),
    "trot_autotune_taylor_workspace": (_sz, [_i, _i, _i, _i]),
    "trot_autotune_taylor": (_i, [_i, _i, _i, _i, _vp, _sz, _vp, _vp, _vp]),
    "trot_set_taylor_tile": (_i, [_i, _i, _i, _i, _i]),
    "trot_get_taylor_tile": (_i, [_i, _i, _i, _i]),
    "trot_force_bias_splitk": (_i, [_vp, _vp, _i, _i, _l, _d, _i, _vp, _vp, _sz, _vp]),
    "trot_field_shift": (_i, [_vp, _vp, _vp, _d, _i, _i, _vp, _vp, _vp, _vp]),
    "trot_vhs_build": (_i, [_vp, _vp, _i, _i, _l, _vp, _vp]),
    "trot_one_body_apply": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "trot_taylor_workspace": (_sz, [_i, _i, _i]),
    "trot_taylor_apply": (_i, [_vp, _vp, _i, _i, _i, _d, _i, _vp, _sz, _vp]),
    "trot_apply_trotter_workspace": (_sz, [_i, _i, _i, _i]),
    "trot_apply_trotter": (_i, [_vp, _i, _vp, _i, _vp, _vp, _vp, _i, _i, _i, _d, _i, _i, _vp, _sz, _vp]),
    "trot_phaseless_update": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _d, _d, _vp]),
    "trot_local_energy_workspace": (_sz, [_i, _i]),
    "trot_local_energy": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _d, _vp, _vp, _sz, _vp]),
    "trot_block_energy": (_i, [_vp, _vp, _vp, _vp, _i, _d, _d, _vp]),
    "trot_rdm1_from_green": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp]),
    "trot_density_corr": (_i, [_vp, _i, _i, _vp, _vp]),
    "trot_weighted_mean_workspace": (_sz, [_i, _l]),
    "trot_weighted_mean": (_i, [_vp, _vp, _vp, _i, _l, _vp, _vp, _sz, _vp]),
    "trot_orthonormalize_workspace": (_sz, [_i, _i, _i]),
    "trot_orthonormalize": (_i, [_vp, _i, _i, _i, _vp, _sz, _vp]),
    "trot_orthogonalize": (_i, [_vp, _i, _i, _i, _vp, _vp, _sz, _vp]),
    "trot_qr_norm_workspace": (_sz, [_i, _i, _i]),
    "trot_qr_norm": (_i, [_vp, _i, _i, _i, _vp, _vp, _sz, _vp]),
    "trot_fp_prepare": (_i, [_vp, _vp, _i, _i, _d, _d, _d, _d, _i, _vp, _vp, _vp, _vp]),
    "trot_scale_rows": (_i, [_vp, _vp, _i, _l, _vp]),
    "trot_fp_weights": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "trot_fp_unnorm": (_i, [_vp, _vp, _vp, _i, _vp]),
    "trot_block_energy_fp": (_i, [_vp, _vp, _vp, _i, _d, _d, _vp, _vp]),
    "trot_sr_workspace": (_sz, [_i]),
    "trot_sr_comb": (_i, [_vp, _d, _i, _vp, _vp, _vp, _sz, _vp]),
    "trot_gather_rows": (_i, [_vp, _vp, _vp, _i, _l, _vp]),
    "trot_sr_comb_dev": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _sz, _vp]),
    "trot_sr_comb_dev_noalloc": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "trot_gemm_batched": (_i, [_i, _i, _i, _i, _i, _i, _i, _d, _d, _vp, _l, _l, _vp, _l, _l, _d, _vp, _l, _l, _i,
                               _vp]),
    "trot_inverse_det": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "trot_split_bf16_problem_bytes": (_sz, []),
    "trot_split_bf16_problem": (_l, [_vp, _l, _i, _i, _vp, _vp]),
    "trot_split_bf16_grouped": (_i, [_vp, _vp, _i, _i, _vp]),
    "trot_gemm_b3_problem_bytes": (_sz, []),
    "trot_gemm_b3_problem": (_l, [_vp, _vp, _i, _i, _i, _vp, _i, _l, _d, _d, _d, _i, _vp, _sz, _vp]),
    "trot_gemm_b3_grouped": (_i, [_i, _vp, _vp, _i, _i, _vp]),
    "trot_b3_reduce_problem_bytes": (_sz, []),
    "trot_b3_reduce_problem": (_l, [_vp, _vp]),
    "trot_b3_reduce_grouped": (_i, [_vp, _vp, _i, _i, _vp]),
    "trot_prop_b3_stage_bytes": (_sz, []),
    "trot_trotter_b3v_stages": (_l, [_vp, _i, _vp, _vp, _vp, _i, _i, _i, _d, _i, _vp, _sz, _i, _vp, _vp, _vp]),
    "trot_prop_b3_grouped": (_i, [_i, _vp, _vp, _i, _i, _vp]),
    "trot_fused_prop_problem_bytes": (_sz, []),
    "trot_fused_prop_problem": (_l, [_vp, _vp, _vp, _vp, _i, _i, _i, _d, _i, _vp]),
    "trot_fused_prop_grouped": (_i, [_vp, _vp, _i, _i, _l, _vp]),
    "trot_gemm_problem_bytes": (_sz, []),
    "trot_gemm_problem": (_l, [_i, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _vp]),
    "trot_gemm_f64_grouped": (_i, [_i, _vp, _vp, _i, _i, _vp]),
    "trot_inv_problem_bytes": (_sz, []),
    "trot_inv_problem": (_l, [_vp, _vp, _vp, _i, _i, _vp]),
    "trot_inv_det_grouped": (_i, [_vp, _vp, _i, _i, _vp]),
    "trot_copy_problem_bytes": (_sz, []),
    "trot_copy_problem": (_i, [_vp, _vp, _l, _vp]),
    "trot_copy_grouped": (_i, [_vp, _i, _l, _vp]),
    "trot_force_bias_slices": (_i, [_i, _i, _l]),
    "trot_force_bias_problem": (_l, [_vp, _vp, _i, _i, _l, _d, _i, _vp, _vp]),
    "trot_sum_problem_bytes": (_sz, []),
    "trot_sum_problem": (_l, [_vp, _vp, _i, _l, _d, _i, _vp]),
    "trot_sum_grouped": (_i, [_vp, _vp, _i, _i, _vp]),
    "trot_shift_problem_bytes": (_sz, []),
    "trot_shift_problem": (_l, [_vp, _vp, _vp, _d, _i, _i, _vp, _vp, _vp, _vp]),
    "trot_shift_grouped": (_i, [_vp, _vp, _i, _i, _vp]),
    "trot_update_problem_bytes": (_sz, []),
    "trot_update_problem": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _d, _d, _vp]),
    "trot_update_grouped": (_i, [_vp, _i, _vp]),
    "trot_cisd_doubles_matrix": (_i, [_vp, _i, _i, _vp, _vp]),
    "trot_pack_sym": (_i, [_vp, _i, _i, _vp, _vp]),
    "trot_cisd_overlap_workspace": (_sz, [_i, _i, _i]),
    "trot_cisd_overlap": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    "trot_cisd_ps": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp]),
    "trot_cisd_gamma": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp]),
    "trot_cisd_energy_workspace": (_sz, [_i, _i, _i, _i, _i, _i, _i]),
    "trot_cisd_energy": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _d, _i, _vp, _vp,
                              _sz, _vp]),
    "trot_trotter_b3v": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _i, _d, _i, _vp, _sz, _vp]),
    "trot_philox_raw": (_i, [_u64, _u64, _u64, _l, _vp, _vp]),
    "trot_philox_normal": (_i, [_vp, _u64, _u64, _l, _vp, _vp]),
    "trot_philox_uniform": (_i, [_vp, _u64, _u64, _vp, _vp]),
    "trot_rng_advance": (_i, [_vp, _u64, _vp]),
    "trot_build_rot_chol": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp]),
    "trot_build_prop_workspace": (_sz, [_i, _i]),
    "trot_build_prop_tensors": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    "trot_vhs_build_lp": (_i, [_vp, _vp, _i, _i, _l, _vp, _i, _i, _vp]),
    "trot_force_bias_lp": (_i, [_vp, _vp, _i, _i, _l, _d, _i, _vp, _i, _vp]),
    "trot_taylor_lp_workspace": (_sz, [_i, _i, _i]),
    "trot_taylor_apply_lp": (_i, [_vp, _vp, _i, _i, _i, _d, _i, _i, _vp, _sz, _vp]),
    "trot_local_energy_lp": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _d, _vp, _i, _vp, _sz, _vp]),
    "trot_symblk_cols": (_l, [_i]),
    "trot_pack_symblk": (_i, [_vp, _i, _i, _vp, _i, _vp]),
    "trot_taylor_apply_symblk": (_i, [_vp, _vp, _i, _i, _i, _d, _i, _vp, _sz, _vp]),
    "trot_taylor_apply_lp_symblk": (_i, [_vp, _vp, _i, _i, _i, _d, _i, _i, _vp, _sz, _vp]),
    "trot_one_body_apply_lp": (_i, [_vp, _vp, _i, _i, _i, _vp, _i, _vp]),
    "trot_taylor_b3_workspace": (_sz, [_i, _i, _i]),
    "trot_taylor_apply_b3_symblk": (_i, [_vp, _vp, _i, _i, _i, _d, _i, _vp, _sz, _vp]),
    "trot_rot_chol_b3_bytes": (_sz, [_i, _i, _i]),
    "trot_rot_chol_b3": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "trot_local_energy_b3_workspace": (_sz, [_i, _i, _i, _i]),
    "trot_local_energy_b3": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _d, _vp, _vp, _sz, _vp]),
    "trot_onebody_b3_planes_bytes": (_sz, [_i]),
    "trot_onebody_b3_planes": (_i, [_vp, _i, _vp, _vp]),
    "trot_trotter_b3_symblk": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _d, _i, _vp, _sz, _vp]),
    "trot_b3_rows_pad": (_l, [_i]),
    "trot_b3_k_pad": (_l, [_i]),
    "trot_b3_planes_bytes": (_sz, [_i, _i, _i]),
    "trot_split_bf16": (_i, [_vp, _i, _i, _i, _l, _i, _i, _vp, _vp]),
    "trot_gemm_b3_workspace": (_sz, [_i, _i, _i, _i]),
    "trot_gemm_b3": (_i, [_vp, _i, _vp, _i, _i, _i, _vp, _i, _l, _d, _d, _d, _i, _vp, _sz, _vp]),
}

PREC = {"f64": 0, "f32": 1, "bf16x3": 2}
F32 = torch.float32
C64 = torch.complex64


def lib():
    """Load libtrot_amd.so (once).  Raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m trot_amd.build` "
                "(trot_amd has no fallback compute path)"
            )
        h = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _check(rc: int, name: str) -> None:
    if rc != TROT_OK:
        msg = lib().trot_last_error().decode(errors="replace")
        raise TrotError(f"{name} failed with status {rc}: {msg}")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _req(t: torch.Tensor, name: str, dtype: torch.dtype, shape: tuple) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: must be a GPU tensor, got {t.device}")
    if t.dtype != dtype:
        raise TypeError(f"{name}: expected {dtype}, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")


def _ws(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


C128 = torch.complex128
F64 = torch.float64


# ---------------------------------------------------------------------------
# wrappers (one per C entry point)
# ---------------------------------------------------------------------------


def overlap_green(trial: torch.Tensor, walkers: torch.Tensor, want_green: bool = True,
                  out_green: torch.Tensor | None = None):
    """(det (W,) c128, green_t (W,N,n) c128 or None) for one spin; ``out_green`` receives green_t when given."""
    W, N, n = walkers.shape
    _req(walkers, "walkers", C128, (W, N, n))
    _req(trial, "trial", F64, (N, n))
    if n > MAX_NOCC_BIG:
        raise ValueError(f"nocc={n} > {MAX_NOCC_BIG} is not supported by the GPU kernels")
    L = lib()
    det = torch.empty((W,), dtype=C128, device=walkers.device)
    gt = None
    if want_green:
        if out_green is not None:
            _req(out_green, "out_green", C128, (W, N, n))
            gt = out_green
        else:
            gt = torch.empty((W, N, n), dtype=C128, device=walkers.device)
    ws = _ws(L.trot_overlap_green_workspace(W, N, n), walkers.device)
    _check(L.trot_overlap_green(_p(trial), _p(walkers), W, N, n, _p(det), _p(gt), _p(ws), ws.numel(), _stream()),
           "trot_overlap_green")
    return det, gt


def overlap_combine(det_a: torch.Tensor, det_b: torch.Tensor | None) -> torch.Tensor:
    W = det_a.shape[0]
    _req(det_a, "det_a", C128, (W,))
    if det_b is not None:
        _req(det_b, "det_b", C128, (W,))
    out = torch.empty((W,), dtype=C128, device=det_a.device)
    _check(lib().trot_overlap_combine(_p(det_a), _p(det_b), _p(out), W, _stream()), "trot_overlap_combine")
    return out


def force_bias(green_t: torch.Tensor, rot_chol_t: torch.Tensor, alpha: float, out: torch.Tensor | None = None,
               accumulate: bool = False) -> torch.Tensor:
    W, N, n = green_t.shape
    G = rot_chol_t.shape[0]
    _req(green_t, "green_t", C128, (W, N, n))
    _req(rot_chol_t, "rot_chol_t", F64, (G, N, n))
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs an output tensor")
        out = torch.empty((W, G), dtype=C128, device=green_t.device)
    _req(out, "fb", C128, (W, G))
    L = lib()
    nbytes = L.trot_force_bias_workspace(W, G, N * n)
    ws = _ws(nbytes, green_t.device) if nbytes else None
    _check(L.trot_force_bias_splitk(_p(green_t), _p(rot_chol_t), W, G, N * n, float(alpha), int(accumulate),
                                    _p(out), _p(ws), 0 if ws is None else ws.numel(), _stream()),
           "trot_force_bias_splitk")
    return out


def force_bias_flat(a: torch.Tensor, r: torch.Tensor, alpha: float) -> torch.Tensor:
    """fb[w, g] = alpha sum_k r[g, k] a[w, k]: a (W, K) c128, r (G, K) f64 (split K for small W)."""
    W, K = a.shape
    G = r.shape[0]
    _req(a, "a", C128, (W, K))
    _req(r, "r", F64, (G, K))
    out = torch.empty((W, G), dtype=C128, device=a.device)
    L = lib()
    nbytes = L.trot_force_bias_workspace(W, G, K)
    ws = _ws(nbytes, a.device) if nbytes else None
    _check(L.trot_force_bias_splitk(_p(a), _p(r), W, G, K, float(alpha), 0, _p(out), _p(ws),
                                    0 if ws is None else ws.numel(), _stream()), "trot_force_bias_splitk")
    return out


def cisd_doubles_matrix(ci2: torch.Tensor) -> torch.Tensor:
    """D[(p,t),(q,u)] = 2 ci2[p,t,q,u] - ci2[p,u,q,t] as an (na nvir, na nvir) f64 matrix."""
    na, va = int(ci2.shape[0]), int(ci2.shape[1])
    _req(ci2, "ci2", F64, (na, va, na, va))
    out = torch.empty((na * va, na * va), dtype=F64, device=ci2.device)
    _check(lib().trot_cisd_doubles_matrix(_p(ci2), na, va, _p(out), _stream()), "trot_cisd_doubles_matrix")
    return out


def pack_sym(m: torch.Tensor) -> torch.Tensor:
    """Upper triangles (row-major) of (count, N, N) symmetric f64 matrices -> (count, N (N+1) / 2)."""
    cnt, N, _ = m.shape
    _req(m, "m", F64, (cnt, N, N))
    out = torch.empty((cnt, N * (N + 1) // 2), dtype=F64, device=m.device)
    _check(lib().trot_pack_sym(_p(m), cnt, N, _p(out), _stream()), "trot_pack_sym")
    return out


def cisd_overlap(gt: torch.Tensor, det: torch.Tensor | None, ci1: torch.Tensor, dmat: torch.Tensor, core: int):
    """(omega, overlap or None, Bs (W, na, nvir)) of the CISD trial from GT = w w[:n]^-1 and det(w[:n])."""
    W, N, n = gt.shape
    na, va = int(ci1.shape[0]), int(ci1.shape[1])
    _req(gt, "green_t", C128, (W, N, n))
    _req(ci1, "ci1", F64, (na, va))
    _req(dmat, "D", F64, (na * va, na * va))
    if det is not None:
        _req(det, "det", C128, (W,))
    omega = torch.empty((W,), dtype=C128, device=gt.device)
    ovl = torch.empty((W,), dtype=C128, device=gt.device) if det is not None else None
    bs = torch.empty((W, na, va), dtype=C128, device=gt.device)
    L = lib()
    ws = _ws(L.trot_cisd_overlap_workspace(W, na, va), gt.device)
    _check(L.trot_cisd_overlap(_p(gt), _p(det), _p(ci1), _p(dmat), W, N, n, core, na, va, _p(omega), _p(ovl), _p(bs),
                               _p(ws), ws.numel(), _stream()), "trot_cisd_overlap")
    return omega, ovl, bs


def cisd_ps(gt: torch.Tensor, bs: torch.Tensor, core: int) -> torch.Tensor:
    W, N, n = gt.shape
    na, va = int(bs.shape[1]), int(bs.shape[2])
    _req(gt, "green_t", C128, (W, N, n))
    _req(bs, "bs", C128, (W, na, va))
    ps = torch.empty((W, N, na), dtype=C128, device=gt.device)
    _check(lib().trot_cisd_ps(_p(gt), _p(bs), W, N, n, core, na, va, _p(ps), _stream()), "trot_cisd_ps")
    return ps


def cisd_gamma(gt: torch.Tensor, ps: torch.Tensor, core: int, packed: bool) -> torch.Tensor:
    """Mixed 1-RDM (W, N, N), or its packed symmetric sums (W, N (N+1) / 2)."""
    W, N, n = gt.shape
    na = int(ps.shape[2])
    _req(gt, "green_t", C128, (W, N, n))
    _req(ps, "ps", C128, (W, N, na))
    shape = (W, N * (N + 1) // 2) if packed else (W, N, N)
    out = torch.empty(shape, dtype=C128, device=gt.device)
    _check(lib().trot_cisd_gamma(_p(gt), _p(ps), W, N, n, core, na, int(packed), _p(out), _stream()),
           "trot_cisd_gamma")
    return out


CISD_ENERGY_WS_BYTES = 1 << 31  # walker chunks sized to keep the energy workspace near 2 GiB


def cisd_energy(gt, bs, ps, omega, lx, dmat, core: int, h0: float, wc: int | None = None,
                d_planes=None) -> torch.Tensor:
    """Local energies; ``d_planes`` (split-BF16 planes of D^T, trot_amd._b3.Planes) puts the doubles term on
    split-BF16 MFMA, else FP64."""
    W, N, n = gt.shape
    na, va = int(bs.shape[1]), int(bs.shape[2])
    G = int(lx.shape[0]) - 1
    _req(gt, "green_t", C128, (W, N, n))
    _req(bs, "bs", C128, (W, na, va))
    _req(ps, "ps", C128, (W, N, na))
    _req(omega, "omega", C128, (W,))
    _req(lx, "lx", F64, (G + 1, N, N))
    _req(dmat, "D", F64, (na * va, na * va))
    L = lib()
    b3 = int(d_planes is not None)
    if b3 and (d_planes.rows != na * va or d_planes.K != na * va or d_planes.complex):
        raise ValueError("d_planes: split-BF16 planes of the real (na nvir)^2 matrix D^T expected")
    if wc is None:
        per = L.trot_cisd_energy_workspace(1, N, n, G, na, va, b3)
        wc = max(1, min(W, CISD_ENERGY_WS_BYTES // max(per, 1)))
    wc = int(min(wc, W))
    e = torch.empty((W,), dtype=C128, device=gt.device)
    ws = _ws(L.trot_cisd_energy_workspace(wc, N, n, G, na, va, b3), gt.device)
    _check(L.trot_cisd_energy(_p(gt), _p(bs), _p(ps), _p(omega), _p(lx), _p(dmat),
                              _p(d_planes.data) if b3 else None, W, N, n, G, core, na, va, float(h0), wc, _p(e),
                              _p(ws), ws.numel(), _stream()), "trot_cisd_energy")
    return e


def field_shift(fields: torch.Tensor, fb: torch.Tensor, mf: torch.Tensor, sqrt_dt: float, out=None):
    """(xbar, shift_term, fb_term); ``out`` = three preallocated tensors of those shapes to write into."""
    W, G = fields.shape
    _req(fields, "fields", F64, (W, G))
    _req(fb, "force_bias", C128, (W, G))
    _req(mf, "mf_shifts", C128, (G,))
    dev = fields.device
    if out is None:
        xbar = torch.empty((W, G), dtype=C128, device=dev)
        shift = torch.empty((W,), dtype=C128, device=dev)
        fbt = torch.empty((W,), dtype=C128, device=dev)
    else:
        xbar, shift, fbt = out
        _req(xbar, "xbar", C128, (W, G))
        _req(shift, "shift_term", C128, (W,))
        _req(fbt, "fb_term", C128, (W,))
    _check(lib().trot_field_shift(_p(fields), _p(fb), _p(mf), float(sqrt_dt), W, G, _p(xbar), _p(shift), _p(fbt),
                                  _stream()), "trot_field_shift")
    return xbar, shift, fbt


_OPS = {"N": 0, "T": 1, "C": 2, "R": 3}


def _mat(x: torch.Tensor, name: str):
    """(batch or None, rows, cols, leading dim, batch stride, complex) of a 2-D or 3-D row-major operand view."""
    if not isinstance(x, torch.Tensor) or x.device.type != "cuda":
        raise ValueError(f"{name}: must be a GPU tensor")
    if x.dtype not in (F64, C128):
        raise TypeError(f"{name}: expected float64 or complex128, got {x.dtype}")
    if x.dim() not in (2, 3):
        raise ValueError(f"{name}: expected a 2-D or 3-D tensor, got shape {tuple(x.shape)}")
    rows, cols = x.shape[-2], x.shape[-1]
    if x.stride(-1) != 1 and cols > 1:
        raise ValueError(f"{name}: rows must be contiguous")
    ld = max(x.stride(-2), cols) if rows > 1 else cols
    if x.stride(-2) < cols and rows > 1:
        raise ValueError(f"{name}: overlapping rows")
    batch = x.shape[0] if x.dim() == 3 else None
    sb = x.stride(0) if (x.dim() == 3 and x.shape[0] > 1) else 0
    return batch, rows, cols, ld, sb, x.is_complex()


def gemm_batched(a: torch.Tensor, b: torch.Tensor, opa: str = "N", opb: str = "N", alpha: complex = 1.0,
                 beta: float = 0.0, out: torch.Tensor | None = None) -> torch.Tensor:
    """out[k] = alpha op(a[k]) op(b[k]) + beta out[k] on FP64 MFMA (trot_gemm_batched).

    op: "N" as stored, "T" transpose, "C" conjugate transpose, "R" conjugate.  a and b are 2-D (shared by
    every batch entry) or 3-D (batch, rows, cols) views with contiguous rows; out (batch, M, N), complex when
    either operand is, may be a strided view (e.g. a column block of a wider matrix).
    """
    ba, ra, ca, lda, sa, ac = _mat(a, "a")
    bb, rb, cb, ldb, sb, bc = _mat(b, "b")
    if ba is not None and bb is not None and ba != bb:
        raise ValueError(f"batch sizes differ: {ba} vs {bb}")
    batch = ba if ba is not None else (bb if bb is not None else 1)
    M, K = (ra, ca) if opa in ("N", "R") else (ca, ra)
    Kb, Nc = (rb, cb) if opb in ("N", "R") else (cb, rb)
    if K != Kb:
        raise ValueError(f"inner dimensions differ: op(a) is {M}x{K}, op(b) is {Kb}x{Nc}")
    cplx = ac or bc
    dt = C128 if cplx else F64
    if out is None:
        if beta != 0.0:
            raise ValueError("beta != 0 needs an output tensor")
        out = torch.empty((batch, M, Nc), dtype=dt, device=a.device)
    if out.dtype != dt or out.device != a.device:
        raise TypeError(f"out: expected {dt} on {a.device}")
    bo, ro, co, ldc, sc, _ = _mat(out, "out")
    if (bo or 1) != batch or ro != M or co != Nc:
        raise ValueError(f"out: expected shape ({batch}, {M}, {Nc}), got {tuple(out.shape)}")
    if batch > 1 and sc == 0:
        raise ValueError("out: needs a batch dimension")
    al = complex(alpha)
    _check(lib().trot_gemm_batched(_OPS[opa], _OPS[opb], int(ac), int(bc), M, Nc, K, al.real, al.imag, _p(a), lda, sa,
                                   _p(b), ldb, sb, float(beta), _p(out), ldc, sc, batch, _stream()),
           "trot_gemm_batched")
    return out


def inverse_det(m: torch.Tensor):
    """(m^-1, det m) of a batch of complex n x n matrices (trot_inverse_det)."""
    W, n, _ = m.shape
    _req(m, "m", C128, (W, n, n))
    minv = torch.empty_like(m)
    det = torch.empty((W,), dtype=C128, device=m.device)
    _check(lib().trot_inverse_det(_p(m), W, n, _p(minv), _p(det), _stream()), "trot_inverse_det")
    return minv, det


def vhs_build(xbar: torch.Tensor, chol_flat: torch.Tensor) -> torch.Tensor:
    W, G = xbar.shape
    N2 = chol_flat.shape[1]
    _req(xbar, "xbar", C128, (W, G))
    _req(chol_flat, "chol_flat", F64, (G, N2))
    out = torch.empty((W, N2), dtype=C128, device=xbar.device)
    _check(lib().trot_vhs_build(_p(xbar), _p(chol_flat), W, G, N2, _p(out), _stream()), "trot_vhs_build")
    return out


def one_body_apply(expm: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    _req(expm, "exp_h1_half", F64, (N, N))
    out = torch.empty_like(w)
    _check(lib().trot_one_body_apply(_p(expm), _p(w), W, N, n, _p(out), _stream()), "trot_one_body_apply")
    return out


def taylor_apply(vhs: torch.Tensor, w: torch.Tensor, sqrt_dt: float, n_terms: int,
                 inplace: bool = False) -> torch.Tensor:
    """sum_k (i sqrt(dt) V)^k/k! w; a new tensor unless ``inplace``."""
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    _req(vhs, "vhs", C128, (W, N, N))
    out = w if inplace else w.clone()
    L = lib()
    ws = _ws(L.trot_taylor_workspace(W, N, n), w.device)
    _check(L.trot_taylor_apply(_p(vhs), _p(out), W, N, n, float(sqrt_dt), int(n_terms), _p(ws), ws.numel(),
                               _stream()), "trot_taylor_apply")
    return out


def apply_trotter(w_a: torch.Tensor, w_b: torch.Tensor | None, xbar: torch.Tensor, chol_flat: torch.Tensor,
                  expm: torch.Tensor, sqrt_dt: float, n_terms: int, chunk: int = 0):
    """New walkers E T_K(i sqrt(dt) V(xbar)) E w for one or two spin sets sharing V."""
    W, N, na = w_a.shape
    G = xbar.shape[1]
    _req(w_a, "walkers", C128, (W, N, na))
    nb = 0
    if w_b is not None:
        nb = w_b.shape[2]
        _req(w_b, "walkers_b", C128, (W, N, nb))
    _req(xbar, "xbar", C128, (W, G))
    _req(chol_flat, "chol_flat", F64, (G, N * N))
    _req(expm, "exp_h1_half", F64, (N, N))
    out_a = w_a.clone()
    out_b = w_b.clone() if w_b is not None else None
    L = lib()
    ws = _ws(L.trot_apply_trotter_workspace(W, N, max(na, nb), int(chunk)), w_a.device)
    _check(L.trot_apply_trotter(_p(out_a), na, _p(out_b), nb, _p(xbar), _p(chol_flat), _p(expm), W, N, G,
                                float(sqrt_dt), int(n_terms), int(chunk), _p(ws), ws.numel(), _stream()),
           "trot_apply_trotter")
    return out_a, out_b


def phaseless_update(weights, ov_old, ov_new, shift, fbt, scalars, nodes, dt, h0_prop, w_floor, w_cap, damping):
    W = weights.shape[0]
    _req(weights, "weights", F64, (W,))
    for t, nm in ((ov_old, "ov_old"), (ov_new, "ov_new"), (shift, "shift_term"), (fbt, "fb_term")):
        _req(t, nm, C128, (W,))
    _req(scalars, "scalars", F64, (2,))
    _req(nodes, "nodes", torch.int64, (1,))
    h0p = complex(h0_prop)
    _check(lib().trot_phaseless_update(_p(weights), _p(ov_old), _p(ov_new), _p(shift), _p(fbt), _p(scalars),
                                       _p(nodes), W, float(dt), h0p.real, h0p.imag, float(w_floor), float(w_cap),
                                       float(damping), _stream()), "trot_phaseless_update")


def local_energy(rot_chol_a, green_t_a, rot_h1t_a, h0: float, rot_chol_b=None, green_t_b=None, rot_h1t_b=None):
    W, N, na = green_t_a.shape
    G = rot_chol_a.shape[0]
    _req(green_t_a, "green_t_a", C128, (W, N, na))
    _req(rot_chol_a, "rot_chol_a", F64, (G, na, N))
    _req(rot_h1t_a, "rot_h1t_a", F64, (N, na))
    nb = 0
    if green_t_b is not None:
        nb = green_t_b.shape[2]
        _req(green_t_b, "green_t_b", C128, (W, N, nb))
        _req(rot_chol_b, "rot_chol_b", F64, (G, nb, N))
        _req(rot_h1t_b, "rot_h1t_b", F64, (N, nb))
    if max(na, nb) > MAX_NOCC_BIG:
        raise ValueError(f"nocc > {MAX_NOCC_BIG} is not supported by the GPU energy kernel")
    L = lib()
    out = torch.empty((W,), dtype=C128, device=green_t_a.device)
    ws = _ws(L.trot_local_energy_workspace(W, G), green_t_a.device)
    _check(L.trot_local_energy(_p(rot_chol_a), _p(rot_chol_b), _p(green_t_a), _p(green_t_b), _p(rot_h1t_a),
                               _p(rot_h1t_b), W, N, na, nb, G, float(h0), _p(out), _p(ws), ws.numel(), _stream()),
           "trot_local_energy")
    return out


def block_energy(energy: torch.Tensor, weights: torch.Tensor, scalars: torch.Tensor, dt: float, shift_ema: float):
    """In place on weights / scalars; returns a (2,) tensor (e_block, sum w)."""
    W = energy.shape[0]
    _req(energy, "energy", C128, (W,))
    _req(weights, "weights", F64, (W,))
    _req(scalars, "scalars", F64, (2,))
    out = torch.empty((2,), dtype=F64, device=energy.device)
    _check(lib().trot_block_energy(_p(energy), _p(weights), _p(scalars), _p(out), W, float(dt), float(shift_ema),
                                   _stream()), "trot_block_energy")
    return out


def rdm1_from_green(green_t: torch.Tensor, trial: torch.Tensor, out: torch.Tensor, spin: int) -> torch.Tensor:
    """out[:, spin] <- green_t @ trial^T for each walker (out: (W, S, N, N) complex128)."""
    W, N, n = green_t.shape
    _req(green_t, "green_t", C128, (W, N, n))
    _req(trial, "trial", F64, (N, n))
    if out.dim() != 4:
        raise ValueError(f"out must be (W, S, N, N), got {tuple(out.shape)}")
    _req(out, "out", C128, (W, out.shape[1], N, N))
    _check(lib().trot_rdm1_from_green(_p(green_t), _p(trial), W, N, n, out.shape[1], int(spin), _p(out), _stream()),
           "trot_rdm1_from_green")
    return out


def density_corr(dm: torch.Tensor) -> torch.Tensor:
    """(W, 3, N, N) complex128 density correlations (uu, ud, dd) from dm (W, 2, N, N)."""
    W, S, N, _ = dm.shape
    _req(dm, "dm", C128, (W, 2, N, N))
    out = torch.empty((W, 3, N, N), dtype=C128, device=dm.device)
    _check(lib().trot_density_corr(_p(dm), W, N, _p(out), _stream()), "trot_density_corr")
    return out


def weighted_mean(x: torch.Tensor, weights: torch.Tensor, wsum: torch.Tensor) -> torch.Tensor:
    """sum_w weights[w] x[w] / wsum (0 when wsum == 0), of shape x.shape[1:]; x float64 or complex128."""
    W = x.shape[0]
    if x.dtype not in (F64, C128) or not x.is_contiguous():
        raise ValueError("x must be a contiguous float64 or complex128 tensor")
    _req(weights, "weights", F64, (W,))
    if wsum.dtype != F64 or wsum.numel() != 1:
        raise ValueError("wsum must be one float64 element")
    L = x.numel() // W * (2 if x.dtype == C128 else 1)
    out = torch.empty(x.shape[1:], dtype=x.dtype, device=x.device)
    lb = lib()
    ws = _ws(lb.trot_weighted_mean_workspace(W, L), x.device)
    _check(lb.trot_weighted_mean(_p(x), _p(weights), _p(wsum), W, L, _p(out), _p(ws), ws.numel(), _stream()),
           "trot_weighted_mean")
    return out


def orthonormalize(w: torch.Tensor) -> torch.Tensor:
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    if n > MAX_NOCC_BIG:
        raise ValueError(f"nocc={n} > {MAX_NOCC_BIG} is not supported")
    out = w.clone()
    L = lib()
    ws = _ws(L.trot_orthonormalize_workspace(W, N, n), w.device)
    _check(L.trot_orthonormalize(_p(out), W, N, n, _p(ws), ws.numel(), _stream()), "trot_orthonormalize")
    return out


def orthogonalize(w: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """(Q, det_r): Q with real positive diag(R) and det_r = prod diag(R) (real, (W,))."""
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    if n > MAX_NOCC_BIG:
        raise ValueError(f"nocc={n} > {MAX_NOCC_BIG} is not supported")
    out = w.clone()
    det_r = torch.empty((W,), dtype=F64, device=w.device)
    L = lib()
    ws = _ws(L.trot_orthonormalize_workspace(W, N, n), w.device)
    _check(L.trot_orthogonalize(_p(out), W, N, n, _p(det_r), _p(ws), ws.numel(), _stream()), "trot_orthogonalize")
    return out, det_r


def qr_norm(w: torch.Tensor) -> torch.Tensor:
    """|det R| of w = QR per walker (W,) FP64, from the Cholesky factor of w^H w."""
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    if n > MAX_NOCC_BIG:
        raise ValueError(f"nocc={n} > {MAX_NOCC_BIG} is not supported")
    det_r = torch.empty((W,), dtype=F64, device=w.device)
    L = lib()
    ws = _ws(L.trot_qr_norm_workspace(W, N, n), w.device)
    _check(L.trot_qr_norm(_p(w), W, N, n, _p(det_r), _p(ws), ws.numel(), _stream()), "trot_qr_norm")
    return det_r


def fp_prepare(fields: torch.Tensor, mf: torch.Tensor, sqrt_dt: float, dt: float, h0p: complex, n_total: int):
    """(xbar (W, G) complex128 = fields + 0i, constants (W,), per-orbital scale constants^(1/n_total) (W,))."""
    W, G = fields.shape
    _req(fields, "fields", F64, (W, G))
    _req(mf, "mf_shifts", C128, (G,))
    xbar = torch.empty((W, G), dtype=C128, device=fields.device)
    cst = torch.empty((W,), dtype=C128, device=fields.device)
    scale = torch.empty((W,), dtype=C128, device=fields.device)
    h0p = complex(h0p)
    _check(lib().trot_fp_prepare(_p(fields), _p(mf), W, G, float(sqrt_dt), float(dt), h0p.real, h0p.imag,
                                 int(n_total), _p(xbar), _p(cst), _p(scale), _stream()), "trot_fp_prepare")
    return xbar, cst, scale


def scale_rows(a: torch.Tensor, scale: torch.Tensor) -> torch.Tensor:
    """In place a[b] *= scale[b] for a complex128 (W, ...) tensor."""
    W = a.shape[0]
    if a.dtype != C128 or not a.is_contiguous() or a.device.type != "cuda":
        raise ValueError("scale_rows: a must be a contiguous complex128 GPU tensor")
    _req(scale, "scale", C128, (W,))
    _check(lib().trot_scale_rows(_p(a), _p(scale), W, a[0].numel() if W else 0, _stream()), "trot_scale_rows")
    return a


def fp_weights(weights: torch.Tensor, det_a: torch.Tensor, det_b: torch.Tensor | None):
    """(norms, weights * norms) with norms = det_a^2 or det_a det_b."""
    W = weights.shape[0]
    _req(weights, "weights", F64, (W,))
    _req(det_a, "det_a", F64, (W,))
    if det_b is not None:
        _req(det_b, "det_b", F64, (W,))
    norms = torch.empty_like(weights)
    w_new = torch.empty_like(weights)
    _check(lib().trot_fp_weights(_p(weights), _p(det_a), _p(det_b), W, _p(norms), _p(w_new), _stream()),
           "trot_fp_weights")
    return norms, w_new


def fp_unnorm(weights: torch.Tensor, norms: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    """In place weights[i] /= norms[idx[i]]."""
    W = weights.shape[0]
    _req(weights, "weights", F64, (W,))
    _req(norms, "norms", F64, (W,))
    _req(idx, "idx", torch.int32, (W,))
    _check(lib().trot_fp_unnorm(_p(weights), _p(norms), _p(idx), W, _stream()), "trot_fp_unnorm")
    return weights


def block_energy_fp(energy: torch.Tensor, weights: torch.Tensor, overlaps: torch.Tensor, dt: float,
                    ene0: float) -> torch.Tensor:
    """(4,) complex128: [energy, weight, overlap, abs_overlap] of the free-projection block."""
    W = energy.shape[0]
    _req(energy, "energy", C128, (W,))
    _req(weights, "weights", F64, (W,))
    _req(overlaps, "overlaps", C128, (W,))
    out = torch.empty((4,), dtype=C128, device=energy.device)
    _check(lib().trot_block_energy_fp(_p(energy), _p(weights), _p(overlaps), W, float(dt), float(ene0), _p(out),
                                      _stream()), "trot_block_energy_fp")
    return out


def sr_comb(weights: torch.Tensor, zeta):
    """Comb indices and weights; ``zeta`` is a float or a one-element FP64 GPU tensor (read on the device)."""
    W = weights.shape[0]
    _req(weights, "weights", F64, (W,))
    L = lib()
    idx = torch.empty((W,), dtype=torch.int32, device=weights.device)
    w_new = torch.empty_like(weights)
    ws = _ws(L.trot_sr_workspace(W), weights.device)
    if isinstance(zeta, torch.Tensor):
        _req(zeta, "zeta", F64, (1,))
        _check(L.trot_sr_comb_dev(_p(weights), _p(zeta), W, _p(idx), _p(w_new), _p(ws), ws.numel(), _stream()),
               "trot_sr_comb_dev")
    else:
        _check(L.trot_sr_comb(_p(weights), float(zeta), W, _p(idx), _p(w_new), _p(ws), ws.numel(), _stream()),
               "trot_sr_comb")
    return idx, w_new


# ---------------------------------------------------------------------------
# counter-based random numbers (Philox4x64-10)
# ---------------------------------------------------------------------------
_U64 = (1 << 64) - 1


def rng_state(seed: int, counter: int, device="cuda") -> torch.Tensor:
    """A device {seed, counter} pair (int64 words holding the uint64 values)."""
    words = [int(seed) & _U64, int(counter) & _U64]
    words = [w - (1 << 64) if w >= (1 << 63) else w for w in words]
    return torch.tensor(words, dtype=torch.int64, device=device)


def _state_args(state, seed, counter):
    if state is not None:
        _req(state, "rng state", torch.int64, (2,))
        return _p(state), 0, 0
    return None, int(seed) & _U64, int(counter) & _U64


def philox_raw(seed: int, counter: int, purpose: int, n_blocks: int, device="cuda") -> torch.Tensor:
    """(n_blocks, 4) raw 64-bit words (as int64) of Philox4x64-10 blocks 0..n_blocks-1."""
    out = torch.empty((n_blocks, 4), dtype=torch.int64, device=device)
    _check(lib().trot_philox_raw(int(seed) & _U64, int(counter) & _U64, int(purpose) & _U64, int(n_blocks), _p(out),
                                 _stream()), "trot_philox_raw")
    return out


def philox_normal(shape, *, state: torch.Tensor | None = None, seed: int = 0, counter: int = 0,
                  device="cuda") -> torch.Tensor:
    """N(0,1) FP64 values of draw ``counter`` (or of the device state's counter when ``state`` is given)."""
    out = torch.empty(shape, dtype=F64, device=device)
    sp, sd, ct = _state_args(state, seed, counter)
    _check(lib().trot_philox_normal(sp, sd, ct, out.numel(), _p(out), _stream()), "trot_philox_normal")
    return out


def philox_uniform(*, state: torch.Tensor | None = None, seed: int = 0, counter: int = 0,
                   device="cuda") -> torch.Tensor:
    """One uniform in [0, 1) as a (1,) FP64 GPU tensor."""
    out = torch.empty((1,), dtype=F64, device=device)
    sp, sd, ct = _state_args(state, seed, counter)
    _check(lib().trot_philox_uniform(sp, sd, ct, _p(out), _stream()), "trot_philox_uniform")
    return out


def rng_advance(state: torch.Tensor, by: int = 1) -> None:
    _req(state, "rng state", torch.int64, (2,))
    _check(lib().trot_rng_advance(_p(state), int(by), _stream()), "trot_rng_advance")


def gather_rows(src: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    W = src.shape[0]
    if not src.is_contiguous() or src.device.type != "cuda":
        raise ValueError("gather_rows: src must be a contiguous GPU tensor")
    _req(idx, "idx", torch.int32, (W,))
    row_bytes = src[0].numel() * src.element_size()
    if row_bytes % 8:
        raise ValueError("gather_rows: rows must be a whole number of doubles")
    dst = torch.empty_like(src)
    _check(lib().trot_gather_rows(_p(src), _p(dst), _p(idx), W, row_bytes // 8, _stream()), "trot_gather_rows")
    return dst


def build_rot_chol(chol: torch.Tensor, trial: torch.Tensor, want_rot: bool = True, want_rot_t: bool = True):
    G, N, _ = chol.shape
    n = trial.shape[1]
    _req(chol, "chol", F64, (G, N, N))
    _req(trial, "trial", F64, (N, n))
    if G > 65535:
        raise ValueError("nchol > 65535 is not supported")
    rc = torch.empty((G, n, N), dtype=F64, device=chol.device) if want_rot else None
    rct = torch.empty((G, N, n), dtype=F64, device=chol.device) if want_rot_t else None
    _check(lib().trot_build_rot_chol(_p(chol), _p(trial), G, N, n, _p(rc), _p(rct), _stream()), "trot_build_rot_chol")
    return rc, rct


def build_prop_tensors(chol: torch.Tensor, dm: torch.Tensor):
    G, N, _ = chol.shape
    _req(chol, "chol", F64, (G, N, N))
    _req(dm, "dm", F64, (N, N))
    dev = chol.device
    L = lib()
    mf = torch.empty((G,), dtype=C128, device=dev)
    mfl = torch.empty((N, N), dtype=C128, device=dev)
    v0m = torch.empty((N, N), dtype=F64, device=dev)
    ws = _ws(L.trot_build_prop_workspace(G, N), dev)
    _check(L.trot_build_prop_tensors(_p(chol), _p(dm), G, N, _p(mf), _p(mfl), _p(v0m), _p(ws), ws.numel(), _stream()),
           "trot_build_prop_tensors")
    return mf, mfl, v0m


# ---------------------------------------------------------------------------
# reduced-precision forms (FP32 / split-BF16 MFMA); FP32 Cholesky-side operands
# ---------------------------------------------------------------------------


def _prec(p: str) -> int:
    if p not in ("f32", "bf16x3"):
        raise ValueError(f"reduced precision must be 'f32' or 'bf16x3', got {p!r}")
    return PREC[p]


def vhs_build_lp(xbar: torch.Tensor, chol_flat32: torch.Tensor, prec: str, out_f64: bool = False) -> torch.Tensor:
    W, G = xbar.shape
    N2 = chol_flat32.shape[1]
    _req(xbar, "xbar", C128, (W, G))
    _req(chol_flat32, "chol_flat (f32)", F32, (G, N2))
    out = torch.empty((W, N2), dtype=C128 if out_f64 else C64, device=xbar.device)
    _check(lib().trot_vhs_build_lp(_p(xbar), _p(chol_flat32), W, G, N2, _p(out), int(out_f64), _prec(prec),
                                   _stream()), "trot_vhs_build_lp")
    return out


def force_bias_lp(green_t: torch.Tensor, rot_chol_t32: torch.Tensor, alpha: float, prec: str,
                  out: torch.Tensor | None = None, accumulate: bool = False) -> torch.Tensor:
    W, N, n = green_t.shape
    G = rot_chol_t32.shape[0]
    _req(green_t, "green_t", C128, (W, N, n))
    _req(rot_chol_t32, "rot_chol_t (f32)", F32, (G, N, n))
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs an output tensor")
        out = torch.empty((W, G), dtype=C128, device=green_t.device)
    _req(out, "fb", C128, (W, G))
    _check(lib().trot_force_bias_lp(_p(green_t), _p(rot_chol_t32), W, G, N * n, float(alpha), int(accumulate),
                                    _p(out), _prec(prec), _stream()), "trot_force_bias_lp")
    return out


def one_body_apply_lp(expm32: torch.Tensor, w: torch.Tensor, prec: str) -> torch.Tensor:
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    _req(expm32, "exp_h1_half (f32)", F32, (N, N))
    out = torch.empty_like(w)
    _check(lib().trot_one_body_apply_lp(_p(expm32), _p(w), W, N, n, _p(out), _prec(prec), _stream()),
           "trot_one_body_apply_lp")
    return out


def taylor_apply_b3_symblk(vpk32: torch.Tensor, w: torch.Tensor, sqrt_dt: float, n_terms: int,
                           inplace: bool = False) -> torch.Tensor:
    """Taylor series with split-BF16 products and a complex128 running sum; V complex64 in sym-block storage."""
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    _req(vpk32, "vhs (sym-block, complex64)", C64, (W, symblk_cols(N)))
    out = w if inplace else w.clone()
    L = lib()
    ws = _ws(L.trot_taylor_b3_workspace(W, N, n), w.device)
    _check(L.trot_taylor_apply_b3_symblk(_p(vpk32), _p(out), W, N, n, float(sqrt_dt), int(n_terms), _p(ws),
                                         ws.numel(), _stream()), "trot_taylor_apply_b3_symblk")
    return out


TAYLOR_TILES = ("64x64", "64x32", "32x64", "32x32")


def autotune_taylor(W: int, N: int, n: int, symblk: bool, device="cuda"):
    """Time the FP64 Taylor-term tile shapes for (W, N, n) and keep the fastest; returns (shape, ms per shape)."""
    L = lib()
    ws = _ws(L.trot_autotune_taylor_workspace(W, N, n, int(symblk)), torch.device(device))
    shape = ctypes.c_int(-1)
    ms = (ctypes.c_float * 4)()
    _check(L.trot_autotune_taylor(W, N, n, int(symblk), _p(ws), ws.numel(), _stream(), ctypes.byref(shape), ms),
           "trot_autotune_taylor")
    return shape.value, [float(x) for x in ms]


def set_taylor_tile(W: int, N: int, n: int, symblk: bool, shape: int) -> None:
    _check(lib().trot_set_taylor_tile(W, N, n, int(symblk), int(shape)), "trot_set_taylor_tile")


def get_taylor_tile(W: int, N: int, n: int, symblk: bool) -> int:
    return int(lib().trot_get_taylor_tile(W, N, n, int(symblk)))


def rot_chol_b3(rot_chol: torch.Tensor) -> torch.Tensor:
    """Split-BF16 planes of rot_chol (G, n, N) for local_energy_b3 (raw bytes, int16 storage)."""
    G, n, N = rot_chol.shape
    _req(rot_chol, "rot_chol", F64, (G, n, N))
    L = lib()
    out = torch.empty(L.trot_rot_chol_b3_bytes(G, n, N) // 2, dtype=torch.int16, device=rot_chol.device)
    _check(L.trot_rot_chol_b3(_p(rot_chol), G, n, N, _p(out), _stream()), "trot_rot_chol_b3")
    return out


def local_energy_b3(rcp_a: torch.Tensor, G: int, green_t_a: torch.Tensor, rot_h1t_a: torch.Tensor, h0: float,
                    rcp_b: torch.Tensor | None = None, green_t_b: torch.Tensor | None = None,
                    rot_h1t_b: torch.Tensor | None = None) -> torch.Tensor:
    """E_L (W,) complex128 with split-BF16 exchange GEMMs; rcp_s from rot_chol_b3."""
    W, N, na = green_t_a.shape
    _req(green_t_a, "green_t_a", C128, (W, N, na))
    _req(rot_h1t_a, "rot_h1t_a", F64, (N, na))
    L = lib()
    nb = 0
    if green_t_b is not None:
        nb = green_t_b.shape[2]
        _req(green_t_b, "green_t_b", C128, (W, N, nb))
        _req(rot_h1t_b, "rot_h1t_b", F64, (N, nb))
    for name, rcp, n in (("a", rcp_a, na), ("b", rcp_b, nb)):
        if rcp is None:
            continue
        if rcp.dtype != torch.int16 or rcp.numel() * 2 != L.trot_rot_chol_b3_bytes(G, n, N):
            raise ValueError(f"local_energy_b3: rcp_{name} must come from rot_chol_b3 for (G, n, N) = ({G}, {n}, {N})")
    out = torch.empty(W, dtype=C128, device=green_t_a.device)
    ws = _ws(L.trot_local_energy_b3_workspace(W, N, G, int(green_t_b is not None)), green_t_a.device)
    _check(L.trot_local_energy_b3(_p(rcp_a), _p(rcp_b), _p(green_t_a), _p(green_t_b), _p(rot_h1t_a),
                                  _p(rot_h1t_b), W, N, na, nb, G, float(h0), _p(out), _p(ws), ws.numel(),
                                  _stream()), "trot_local_energy_b3")
    return out


def onebody_b3_planes(expm: torch.Tensor) -> torch.Tensor:
    """Split bf16 planes of (expm - I) for trotter_b3_symblk (raw bytes, int16 storage)."""
    N = expm.shape[0]
    _req(expm, "expm", F64, (N, N))
    L = lib()
    out = torch.empty(L.trot_onebody_b3_planes_bytes(N) // 2, dtype=torch.int16, device=expm.device)
    _check(L.trot_onebody_b3_planes(_p(expm), N, _p(out), _stream()), "trot_onebody_b3_planes")
    return out


def trotter_b3_symblk(vpk32: torch.Tensor, w: torch.Tensor, onebody_planes: torch.Tensor, sqrt_dt: float,
                      n_terms: int, out: torch.Tensor | None = None) -> torch.Tensor:
    """E . taylor(i sqrt(dt) V, E . w, K) on split-BF16 MFMA; ``out`` may be ``w`` (in place)."""
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    _req(vpk32, "vhs (sym-block, complex64)", C64, (W, symblk_cols(N)))
    L = lib()
    if onebody_planes.dtype != torch.int16 or onebody_planes.numel() * 2 != L.trot_onebody_b3_planes_bytes(N):
        raise ValueError("trotter_b3_symblk: onebody_planes must come from onebody_b3_planes for this N")
    if out is None:
        out = torch.empty_like(w)
    else:
        _req(out, "out", C128, (W, N, n))
    ws = _ws(L.trot_taylor_b3_workspace(W, N, n), w.device)
    _check(L.trot_trotter_b3_symblk(_p(vpk32), _p(onebody_planes), _p(w), _p(out), W, N, n, float(sqrt_dt),
                                    int(n_terms), _p(ws), ws.numel(), _stream()), "trot_trotter_b3_symblk")
    return out


def trotter_b3v(vsplit: torch.Tensor, w: torch.Tensor, onebody_planes: torch.Tensor, sqrt_dt: float, n_terms: int,
                hi_from: int, out: torch.Tensor | None = None) -> torch.Tensor:
    """trotter_b3_symblk with V as split planes (int32 (2, W, symblk_cols(N)), from _b3.gemm_nt_split);
    Taylor terms k >= hi_from are one bf16 pass on the hi planes of V and of the term."""
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    _req(vsplit, "vhs (split planes)", torch.int32, (2, W, symblk_cols(N)))
    L = lib()
    if onebody_planes.dtype != torch.int16 or onebody_planes.numel() * 2 != L.trot_onebody_b3_planes_bytes(N):
        raise ValueError("trotter_b3v: onebody_planes must come from onebody_b3_planes for this N")
    if out is None:
        out = torch.empty_like(w)
    else:
        _req(out, "out", C128, (W, N, n))
    ws = _ws(L.trot_taylor_b3_workspace(W, N, n), w.device)
    _check(L.trot_trotter_b3v(_p(vsplit), int(hi_from), _p(onebody_planes), _p(w), _p(out), W, N, n, float(sqrt_dt),
                              int(n_terms), _p(ws), ws.numel(), _stream()), "trot_trotter_b3v")
    return out


def trotter_b3v_stages(vsplit: torch.Tensor, w_in: torch.Tensor, onebody_planes: torch.Tensor, sqrt_dt: float,
                       n_terms: int, hi_from: int, w_out: torch.Tensor, ws: torch.Tensor):
    """The launches trotter_b3v would make (trot_trotter_b3v_stages): a list of (kind, workgroups, descriptor
    bytes); ``ws`` (uint8, trot_taylor_b3_workspace bytes) must outlive their use."""
    W, N, n = w_in.shape
    _req(w_in, "walkers", C128, (W, N, n))
    _req(w_out, "out", C128, (W, N, n))
    _req(vsplit, "vhs (split planes)", torch.int32, (2, W, symblk_cols(N)))
    L = lib()
    if onebody_planes.dtype != torch.int16 or onebody_planes.numel() * 2 != L.trot_onebody_b3_planes_bytes(N):
        raise ValueError("trotter_b3v_stages: onebody_planes must come from onebody_b3_planes for this N")
    size = L.trot_prop_b3_stage_bytes()
    mx = 4 + (W + 7) // 8 * max(int(n_terms), 1)  # split, two one-body factors, terms of chunks of >= 8 walkers
    kinds = (ctypes.c_int * mx)()
    wgs = (ctypes.c_long * mx)()
    descs = ctypes.create_string_buffer(size * mx)
    r = L.trot_trotter_b3v_stages(_p(vsplit), int(hi_from), _p(onebody_planes), _p(w_in), _p(w_out), W, N, n,
                                  float(sqrt_dt), int(n_terms), _p(ws), ws.numel(), mx, kinds, wgs, descs)
    _check(0 if r >= 0 else int(r), "trot_trotter_b3v_stages")
    raw = descs.raw
    return [(int(kinds[i]), int(wgs[i]), raw[i * size:(i + 1) * size]) for i in range(r)]


def taylor_apply_lp(vhs32: torch.Tensor, w: torch.Tensor, sqrt_dt: float, n_terms: int, prec: str,
                    inplace: bool = False) -> torch.Tensor:
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    _req(vhs32, "vhs (complex64)", C64, (W, N, N))
    out = w if inplace else w.clone()
    L = lib()
    ws = _ws(L.trot_taylor_lp_workspace(W, N, n), w.device)
    _check(L.trot_taylor_apply_lp(_p(vhs32), _p(out), W, N, n, float(sqrt_dt), int(n_terms), _prec(prec), _p(ws),
                                  ws.numel(), _stream()), "trot_taylor_apply_lp")
    return out


def local_energy_lp(rot_chol_a32, green_t_a, rot_h1t_a, h0: float, prec: str, rot_chol_b32=None, green_t_b=None,
                    rot_h1t_b=None):
    W, N, na = green_t_a.shape
    G = rot_chol_a32.shape[0]
    _req(green_t_a, "green_t_a", C128, (W, N, na))
    _req(rot_chol_a32, "rot_chol_a (f32)", F32, (G, na, N))
    _req(rot_h1t_a, "rot_h1t_a", F64, (N, na))
    nb = 0
    if green_t_b is not None:
        nb = green_t_b.shape[2]
        _req(green_t_b, "green_t_b", C128, (W, N, nb))
        _req(rot_chol_b32, "rot_chol_b (f32)", F32, (G, nb, N))
        _req(rot_h1t_b, "rot_h1t_b", F64, (N, nb))
    if max(na, nb) > MAX_NOCC_BIG:
        raise ValueError(f"nocc > {MAX_NOCC_BIG} is not supported by the GPU energy kernel")
    L = lib()
    out = torch.empty((W,), dtype=C128, device=green_t_a.device)
    ws = _ws(L.trot_local_energy_workspace(W, G), green_t_a.device)
    _check(L.trot_local_energy_lp(_p(rot_chol_a32), _p(rot_chol_b32), _p(green_t_a), _p(green_t_b), _p(rot_h1t_a),
                                  _p(rot_h1t_b), W, N, na, nb, G, float(h0), _p(out), _prec(prec), _p(ws),
                                  ws.numel(), _stream()), "trot_local_energy_lp")
    return out


# ---------------------------------------------------------------------------
# symmetric Cholesky vectors: sym-block storage of L_g and V
# ---------------------------------------------------------------------------


def symblk_cols(N: int) -> int:
    """Numbers per N x N symmetric matrix in sym-block storage (upper 8x8 blocks)."""
    return int(lib().trot_symblk_cols(int(N)))


def pack_symblk(chol: torch.Tensor, out_f32: bool = False) -> torch.Tensor:
    G, N, _ = chol.shape
    _req(chol, "chol", F64, (G, N, N))
    if G > 65535:
        raise ValueError("nchol > 65535 is not supported")
    out = torch.empty((G, symblk_cols(N)), dtype=F32 if out_f32 else F64, device=chol.device)
    _check(lib().trot_pack_symblk(_p(chol), G, N, _p(out), int(out_f32), _stream()), "trot_pack_symblk")
    return out


def taylor_apply_symblk(vpk: torch.Tensor, w: torch.Tensor, sqrt_dt: float, n_terms: int,
                        inplace: bool = False) -> torch.Tensor:
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    _req(vpk, "vhs (sym-block)", C128, (W, symblk_cols(N)))
    out = w if inplace else w.clone()
    L = lib()
    ws = _ws(L.trot_taylor_workspace(W, N, n), w.device)
    _check(L.trot_taylor_apply_symblk(_p(vpk), _p(out), W, N, n, float(sqrt_dt), int(n_terms), _p(ws), ws.numel(),
                                      _stream()), "trot_taylor_apply_symblk")
    return out


def taylor_apply_lp_symblk(vpk32: torch.Tensor, w: torch.Tensor, sqrt_dt: float, n_terms: int, prec: str,
                           inplace: bool = False) -> torch.Tensor:
    W, N, n = w.shape
    _req(w, "walkers", C128, (W, N, n))
    _req(vpk32, "vhs (sym-block, complex64)", C64, (W, symblk_cols(N)))
    out = w if inplace else w.clone()
    L = lib()
    ws = _ws(L.trot_taylor_lp_workspace(W, N, n), w.device)
    _check(L.trot_taylor_apply_lp_symblk(_p(vpk32), _p(out), W, N, n, float(sqrt_dt), int(n_terms), _prec(prec),
                                         _p(ws), ws.numel(), _stream()), "trot_taylor_apply_lp_symblk")
    return out


# ---------------------------------------------------------------------------
# grouped launches over many walker sets
# ---------------------------------------------------------------------------
GOP = {"one_body": 0, "taylor": 1, "overlap": 2, "green": 3, "vhs": 4, "taylor_narrow": 5, "force_bias": 6,
       "taylor_64x64": 7, "taylor_64x32": 8, "one_body_narrow": 9}
TROT_B3_STAGE_COPY = 5  # a propagation launch kind without a grouped form (include/trot_amd.h)
# trot_autotune_taylor's shapes (0 = 64x64, 1 = 64x32, 2 = 32x64, 3 = 32x32) as grouped Taylor operations
TAYLOR_GOP_OF_SHAPE = {0: "taylor_64x64", 1: "taylor_64x32", 2: "taylor", 3: "taylor_narrow"}


def fused_prop_lds(N: int, n: int) -> int:
    """Bytes of LDS the fused propagation of a walker set takes: its N x n operand, padded to 16, complex FP64."""
    return ((N + 15) // 16 * 16) * ((n + 15) // 16 * 16) * 16


class Grouped:
    """Descriptors of one operation over many walker sets, gathered into device memory, and its launch.

    ``add(...)`` takes the operands of one set (tensors whose storage must outlive every launch; the
    descriptors hold their addresses); ``finish()`` uploads the table; ``launch()`` runs all sets in one
    kernel launch on the current stream."""

    def __init__(self, kind: str, op: str | None = None):
        self.kind, self.op = kind, op
        L = lib()
        self.size = {"gemm": L.trot_gemm_problem_bytes, "inv": L.trot_inv_problem_bytes,
                     "copy": L.trot_copy_problem_bytes, "sum": L.trot_sum_problem_bytes,
                     "shift": L.trot_shift_problem_bytes, "update": L.trot_update_problem_bytes,
                     "fused": L.trot_fused_prop_problem_bytes, "b3stage": L.trot_prop_b3_stage_bytes,
                     "b3split": L.trot_split_bf16_problem_bytes, "b3gemm": L.trot_gemm_b3_problem_bytes,
                     "b3reduce": L.trot_b3_reduce_problem_bytes}[kind]()
        self.max_lds = 0
        self.descs, self.counts, self.keep = [], [], []
        self.max_bytes = 0
        self.probs = self.start = None

    def add_gemm(self, a, b, c, d, W: int, N: int, n: int, coef: float = 0.0):
        buf = ctypes.create_string_buffer(self.size)
        wg = lib().trot_gemm_problem(GOP[self.op], _p(a), _p(b), _p(c), _p(d), int(W), int(N), int(n), float(coef),
                                     buf)
        _check(0 if wg > 0 else int(wg), "trot_gemm_problem")
        self.descs.append(buf.raw)
        self.counts.append(int(wg))
        self.keep += [a, b, c, d]

    def add_inv(self, m, minv, det, W: int, n: int):
        buf = ctypes.create_string_buffer(self.size)
        r = lib().trot_inv_problem(_p(m), _p(minv), _p(det), int(W), int(n), buf)
        _check(0 if r > 0 else int(r), "trot_inv_problem")
        self.descs.append(buf.raw)
        self.counts.append(int(r))
        self.keep += [m, minv, det]

    def _add(self, buf, count: int, keep):
        self.descs.append(buf.raw)
        self.counts.append(int(count))
        self.keep += list(keep)

    def add_force_bias(self, green_t, rot_chol_t, W: int, G: int, Nn: int, alpha: float, accumulate: bool, out):
        """trot_force_bias_splitk's GEMM: ``out`` is fb with one slice, else the partial products."""
        buf = ctypes.create_string_buffer(self.size)
        wg = lib().trot_force_bias_problem(_p(green_t), _p(rot_chol_t), int(W), int(G), int(Nn), float(alpha),
                                           int(accumulate), _p(out), buf)
        _check(0 if wg > 0 else int(wg), "trot_force_bias_problem")
        self._add(buf, wg, (green_t, rot_chol_t, out))

    def add_sum(self, src, dst, nb: int, length: int, scale: float, accumulate: bool):
        buf = ctypes.create_string_buffer(self.size)
        nblk = lib().trot_sum_problem(_p(src), _p(dst), int(nb), int(length), float(scale), int(accumulate), buf)
        _check(0 if nblk > 0 else int(nblk), "trot_sum_problem")
        self._add(buf, nblk, (src, dst))

    def add_shift(self, fields, fb, mf, sqrt_dt: float, W: int, G: int, xbar, shift, fbt):
        buf = ctypes.create_string_buffer(self.size)
        r = lib().trot_shift_problem(_p(fields), _p(fb), _p(mf), float(sqrt_dt), int(W), int(G), _p(xbar), _p(shift),
                                     _p(fbt), buf)
        _check(0 if r > 0 else int(r), "trot_shift_problem")
        self._add(buf, r, (fields, fb, mf, xbar, shift, fbt))

    def add_update(self, det_a, det_b, overlaps, shift, fbt, weights, scal, nodes, W: int, dt: float, h0p: complex,
                   w_floor: float, w_cap: float, damping: float):
        buf = ctypes.create_string_buffer(self.size)
        h0p = complex(h0p)
        _check(lib().trot_update_problem(_p(det_a), _p(det_b), _p(overlaps), _p(shift), _p(fbt), _p(weights),
                                         _p(scal), _p(nodes), int(W), float(dt), h0p.real, h0p.imag, float(w_floor),
                                         float(w_cap), float(damping), buf), "trot_update_problem")
        self._add(buf, 1, (det_a, det_b, overlaps, shift, fbt, weights, scal, nodes))

    def add_fused(self, vsym, expm, w, acc, W: int, N: int, n: int, sqrt_dt: float, n_terms: int):
        """The whole FP64 propagation E . T_K . E of one walker set (trot_fused_prop_problem), or False when the
        set is too large for its LDS-resident series."""
        buf = ctypes.create_string_buffer(self.size)
        r = lib().trot_fused_prop_problem(_p(vsym), _p(expm), _p(w), _p(acc), int(W), int(N), int(n), float(sqrt_dt),
                                          int(n_terms), buf)
        if r == -2:
            return False
        _check(0 if r > 0 else int(r), "trot_fused_prop_problem")
        self._add(buf, r, (vsym, expm, w, acc))
        self.max_lds = max(self.max_lds, fused_prop_lds(N, n))
        return True

    def add_b3_stage(self, desc: bytes, workgroups: int, keep=()):
        """One set's descriptor of a propagation stage (from trotter_b3v_stages), for kind ``self.op``."""
        if len(desc) != self.size:
            raise ValueError("add_b3_stage: descriptor size")
        self.descs.append(desc)
        self.counts.append(int(workgroups))
        self.keep += list(keep)

    def add_b3_split(self, src: torch.Tensor, dst: torch.Tensor):
        """Split-BF16 planes of a complex128 row-major matrix (trot_split_bf16_problem) into ``dst`` (uint8)."""
        rows, K = src.shape
        _req(src, "src", C128, (rows, K))
        if dst.dtype != torch.uint8 or dst.numel() < lib().trot_b3_planes_bytes(rows, K, 1):
            raise ValueError("add_b3_split: dst too small")
        buf = ctypes.create_string_buffer(self.size)
        r = lib().trot_split_bf16_problem(_p(src), K, rows, K, _p(dst), buf)
        _check(0 if r > 0 else int(r), "trot_split_bf16_problem")
        self._add(buf, r, (src, dst))

    def add_b3_gemm(self, a_data: torch.Tensor, b_data: torch.Tensor, M: int, N: int, K: int, out: torch.Tensor,
                    ldc: int, alpha: complex, beta: float, ksplit: int, ws: torch.Tensor | None):
        """One complex-A product of trot_gemm_b3 (kind ``self.op``); returns its descriptor for a reduction."""
        buf = ctypes.create_string_buffer(self.size)
        al = complex(alpha)
        r = lib().trot_gemm_b3_problem(_p(a_data), _p(b_data), int(M), int(N), int(K), _p(out), int(self.op), int(ldc),
                                       al.real, al.imag, float(beta), int(ksplit), _p(ws),
                                       0 if ws is None else ws.numel(), buf)
        _check(0 if r > 0 else int(r), "trot_gemm_b3_problem")
        self._add(buf, r, (a_data, b_data, out) + (() if ws is None else (ws,)))
        return buf.raw

    def add_b3_reduce(self, gemm_desc: bytes) -> bool:
        """The slice reduction of a split-K product from add_b3_gemm; False when it has no split K."""
        buf = ctypes.create_string_buffer(self.size)
        r = lib().trot_b3_reduce_problem(gemm_desc, buf)
        _check(0 if r >= 0 else int(r), "trot_b3_reduce_problem")
        if r == 0:
            return False
        self._add(buf, r, ())
        return True

    def add_copy(self, src: torch.Tensor, dst: torch.Tensor):
        if src.numel() != dst.numel() or src.dtype != dst.dtype or not (src.is_contiguous() and dst.is_contiguous()):
            raise ValueError("copy: matching contiguous tensors needed")
        nbytes = src.numel() * src.element_size()
        buf = ctypes.create_string_buffer(self.size)
        _check(lib().trot_copy_problem(_p(src), _p(dst), nbytes, buf), "trot_copy_problem")
        self.descs.append(buf.raw)
        self.counts.append(1)
        self.max_bytes = max(self.max_bytes, nbytes)
        self.keep += [src, dst]

    def finish(self, device="cuda"):
        raw = b"".join(self.descs)
        self.probs = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)
        starts = [0]
        for c in self.counts:
            starts.append(starts[-1] + c)
        if starts[-1] >= 2**31:
            raise ValueError("grouped launch too large")
        self.start = torch.tensor(starts, dtype=torch.int32, device=device)
        self.total = starts[-1]
        return self

    def launch(self):
        L = lib()
        n = len(self.descs)
        if self.kind == "gemm":
            _check(L.trot_gemm_f64_grouped(GOP[self.op], _p(self.probs), _p(self.start), n, self.total, _stream()),
                   "trot_gemm_f64_grouped")
        elif self.kind == "inv":
            _check(L.trot_inv_det_grouped(_p(self.probs), _p(self.start), n, self.total, _stream()),
                   "trot_inv_det_grouped")
        elif self.kind == "sum":
            _check(L.trot_sum_grouped(_p(self.probs), _p(self.start), n, self.total, _stream()), "trot_sum_grouped")
        elif self.kind == "shift":
            _check(L.trot_shift_grouped(_p(self.probs), _p(self.start), n, self.total, _stream()),
                   "trot_shift_grouped")
        elif self.kind == "update":
            _check(L.trot_update_grouped(_p(self.probs), n, _stream()), "trot_update_grouped")
        elif self.kind == "b3split":
            _check(L.trot_split_bf16_grouped(_p(self.probs), _p(self.start), n, self.total, _stream()),
                   "trot_split_bf16_grouped")
        elif self.kind == "b3gemm":
            _check(L.trot_gemm_b3_grouped(int(self.op), _p(self.probs), _p(self.start), n, self.total, _stream()),
                   "trot_gemm_b3_grouped")
        elif self.kind == "b3reduce":
            _check(L.trot_b3_reduce_grouped(_p(self.probs), _p(self.start), n, self.total, _stream()),
                   "trot_b3_reduce_grouped")
        elif self.kind == "b3stage":
            _check(L.trot_prop_b3_grouped(int(self.op), _p(self.probs), _p(self.start), n, self.total, _stream()),
                   "trot_prop_b3_grouped")
        elif self.kind == "fused":
            _check(L.trot_fused_prop_grouped(_p(self.probs), _p(self.start), n, self.total, self.max_lds, _stream()),
                   "trot_fused_prop_grouped")
        else:
            _check(L.trot_copy_grouped(_p(self.probs), n, self.max_bytes, _stream()), "trot_copy_grouped")
