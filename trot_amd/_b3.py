"""Split-BF16 operands and the 3-pass split-BF16 NT GEMM (include/trot_amd.h, trot_gemm_b3).

A ``Planes`` object holds one operand matrix (rows x K) as bf16 hi / lo
planes on the GPU, in the padded K-contiguous storage the kernel reads with
16-byte vector loads.  Cholesky-side operands are split once per context;
walker-side operands (shifted fields, Green's functions) once per use.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from ._lib import C128, C64, F32, F64, _check, _p, _stream, _ws, lib


@dataclass
class Planes:
    data: torch.Tensor  # uint8 storage of the bf16 planes
    rows: int
    K: int
    complex: bool


def split(src: torch.Tensor, K: int | None = None, rows: int | None = None, transposed: bool = False) -> Planes:
    """Split a 2-D FP64/FP32, real or complex GPU tensor (rows x K; or its transpose) into bf16 planes."""
    if src.device.type != "cuda" or not src.is_contiguous() or src.dim() != 2:
        raise ValueError("split: need a contiguous 2-D GPU tensor")
    cx = src.is_complex()
    if src.dtype not in (F64, F32, C128, C64):
        raise TypeError(f"split: unsupported dtype {src.dtype}")
    f32 = src.dtype in (F32, C64)
    if transposed:
        K_, rows_ = src.shape
    else:
        rows_, K_ = src.shape
    ld = src.shape[1]
    L = lib()
    nbytes = L.trot_b3_planes_bytes(rows_, K_, int(cx))
    out = torch.empty(nbytes, dtype=torch.uint8, device=src.device)
    _check(L.trot_split_bf16(_p(src), int(f32), int(cx), int(transposed), ld, rows_, K_, _p(out), _stream()),
           "trot_split_bf16")
    return Planes(out, rows_, K_, cx)


def gemm_nt(a: Planes, b: Planes, out: torch.Tensor, alpha: complex = 1.0, beta: float = 0.0,
            ksplit: int = 1) -> torch.Tensor:
    """out (M x N) = alpha a b^T (+ beta out); out complex64/complex128 (complex a) or float32/float64."""
    if b.complex:
        raise ValueError("gemm_nt: B must be real")
    if a.K != b.K:
        raise ValueError(f"gemm_nt: K mismatch {a.K} != {b.K}")
    M, N = a.rows, b.rows
    if out.dim() != 2 or out.shape[0] != M or out.shape[1] < N or not out.is_contiguous():
        raise ValueError(f"gemm_nt: output must be contiguous ({M}, >= {N}), got {tuple(out.shape)}")
    want = (C128, C64) if a.complex else (F64, F32)
    if out.dtype not in want:
        raise TypeError(f"gemm_nt: output dtype {out.dtype} not in {want}")
    c_f64 = out.dtype in (C128, F64)
    alpha = complex(alpha)
    L = lib()
    ws = _ws(L.trot_gemm_b3_workspace(M, N, int(a.complex), int(ksplit)), out.device)
    _check(L.trot_gemm_b3(_p(a.data), int(a.complex), _p(b.data), M, N, a.K, _p(out), int(c_f64), out.shape[1],
                          alpha.real, alpha.imag, float(beta), int(ksplit), _p(ws), ws.numel(), _stream()),
           "trot_gemm_b3")
    return out


def gemm_nt_split(a: Planes, b: Planes, out: torch.Tensor | None = None) -> torch.Tensor:
    """V = a b^T for complex a, written as split complex64: an int32 (2, M, N) tensor of words
    bf16(re) | bf16(im) << 16, hi plane then lo plane (trot_gemm_b3 output kind 2)."""
    if b.complex or not a.complex:
        raise ValueError("gemm_nt_split: A complex, B real")
    if a.K != b.K:
        raise ValueError(f"gemm_nt_split: K mismatch {a.K} != {b.K}")
    M, N = a.rows, b.rows
    if out is None:
        out = torch.empty((2, M, N), dtype=torch.int32, device=a.data.device)
    elif out.dtype != torch.int32 or tuple(out.shape) != (2, M, N) or not out.is_contiguous():
        raise ValueError(f"gemm_nt_split: out must be a contiguous int32 (2, {M}, {N}) tensor")
    L = lib()
    _check(L.trot_gemm_b3(_p(a.data), 1, _p(b.data), M, N, a.K, _p(out), 2, N, 1.0, 0.0, 0.0, 1, None, 0,
                          _stream()), "trot_gemm_b3")
    return out
