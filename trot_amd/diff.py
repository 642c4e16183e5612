"""d E_block / d lambda for a one-body perturbation h1 + lambda D, by tangent (forward-mode) propagation.

The block energy after K phaseless steps (trot/prop/afqmc.py:82-144, trot/prop/blocks.py:236-257) depends
on lambda through the one-body propagator E = exp(-dt/2 h1_eff) (trot/prop/chol_afqmc_ops.py:90-131) and
through the one-body energy term (trot/meas/rhf.py:128-138).  With one scalar parameter the derivative is
carried forward next to the FP64 step, every quantity x with its tangent dx = dx/dlambda:

* Green's function GT = w m^-1 (m = C^T w): dGT = (dw - GT C^T dw) m^-1;  dlog(overlap) = 2 tr(m^-1 C^T dw);
* force bias and VHS are linear: dfb = fb(dGT) and dV = V(dxbar) on the same HIP kernels;
* the one-body factors and the Taylor series act on [dw; w] through the block-triangular matrices
  [[E, dE], [0, E]] and i sqrt(dt) [[V, dV], [0, V]]: their powers carry the product-rule terms, so the
  HIP one-body and Taylor kernels (at 2N) produce w and dw in one pass;  dE is the Frechet derivative of
  the exponential, the top-right block of expm([[A, B], [0, A]]) with A = -dt/2 h1_eff, B = -dt/2 D;
* the phaseless weight update is differentiated through |I| cos(theta) with the reference's masks
  (masked importances and capped weights carry no derivative);
* E_L is h0 plus a linear and a quadratic form in GT, so its tangent is exactly
  [E_L(GT + s dGT) - E_L(GT - s dGT)] / (2 s) (s only balances rounding) plus the explicit 2 sum GT o (D C).

Forward mode needs no stored trajectory: memory does not grow with the number of steps, and the cost is
about three primal steps per step.  ``checkpoint_every`` keeps the full (primal, tangent) state every k
steps; ``resume`` recomputes from any of them.  ``BlockEnergy`` exposes the result through a
torch.autograd.Function, so ``E.backward()`` gives dE/dlambda in ``lam.grad``.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Callable, List

import numpy as np
import scipy.linalg
import torch

from . import _lib, meas, prop, trial as trial_mod
from .core import HamChol, System

C128 = torch.complex128


@dataclass
class TangentState:
    """Primal state of the step and its derivative with respect to lambda (restricted walkers)."""

    w: torch.Tensor  # (W, N, n) complex128
    dw: torch.Tensor
    weights: torch.Tensor  # (W,) float64
    dweights: torch.Tensor
    overlaps: torch.Tensor  # (W,) complex128
    dlogov: torch.Tensor  # d log(overlap)
    scal: torch.Tensor  # [pop_control_ene_shift, e_estimate]
    dscal: torch.Tensor
    nodes: torch.Tensor
    step: int = 0

    def clone(self) -> "TangentState":
        return replace(self, **{k: getattr(self, k).clone() for k in
                                ("w", "dw", "weights", "dweights", "overlaps", "dlogov", "scal", "dscal", "nodes")})


@dataclass
class TangentCtx:
    """Everything a tangent step needs at one lambda."""

    N: int
    n: int
    dt: float
    sqrt_dt: float
    n_terms: int
    h0: float
    h0_prop: complex
    mf: torch.Tensor  # (G,) complex128
    chol_flat: torch.Tensor  # (G, N*N)
    E_aug: torch.Tensor  # (2N, 2N) float64: [[E, dE], [0, E]]
    C: torch.Tensor  # (N, n) float64
    rot_chol: torch.Tensor  # (G, n, N)
    rot_chol_t: torch.Tensor  # (G, N, n)
    rot_h1t: torch.Tensor  # (N, n)
    d_rot_h1t: torch.Tensor  # D C (N, n)
    params: prop.QmcParams


def build_ctx(h0: float, h1: np.ndarray, chol: torch.Tensor, c: np.ndarray, D: np.ndarray, lam: float,
              params: prop.QmcParams) -> TangentCtx:
    """Propagation and measurement contexts of h1 + lam D (the product's own builders), and dE/dlam."""
    h1l = np.asarray(h1, dtype=np.float64) + lam * np.asarray(D, dtype=np.float64)
    N, n = c.shape
    ham = HamChol(h0, h1l, chol).to_device()
    sys_ = System(N, (n, n), "restricted")
    tr = trial_mod.RhfTrial(c)
    tops = trial_mod.make_rhf_trial_ops(sys_)
    pctx = prop.make_prop_ops("restricted", "restricted").build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = meas.make_rhf_meas_ops(sys_).build_meas_ctx(ham, tr)
    # h1_eff = E's generator: -2/dt log E is avoided by rebuilding it from its definition
    # (trot/prop/chol_afqmc_ops.py:90-100): h1 - 1/2 sum_g L_g L_g - sum_g Re(i mf_g) L_g
    mf = pctx.mf_shifts
    _, mfl, v0m = _lib.build_prop_tensors(ham.chol, prop._total_dm(tops.get_rdm1(tr)))
    h1_eff = torch.as_tensor(h1l, device="cuda") - v0m - (1j * mfl).real
    A = (-0.5 * params.dt) * h1_eff.cpu().numpy()
    B = (-0.5 * params.dt) * np.asarray(D, dtype=np.float64)
    big = scipy.linalg.expm(np.block([[A, B], [np.zeros_like(A), A]]))
    E = pctx.exp_h1_half.cpu().numpy()  # the primal step's own E
    E_aug = np.block([[E, big[:N, N:]], [np.zeros_like(E), E]])
    Ct = torch.as_tensor(c, dtype=torch.float64, device="cuda")
    sp = mctx.spin
    return TangentCtx(N=N, n=n, dt=pctx.dt, sqrt_dt=pctx.sqrt_dt, n_terms=params.n_exp_terms, h0=float(h0),
                      h0_prop=complex(pctx.h0_prop), mf=mf, chol_flat=pctx.chol_flat,
                      E_aug=torch.as_tensor(E_aug, device="cuda").contiguous(), C=Ct, rot_chol=sp.rot_chol,
                      rot_chol_t=sp.rot_chol_t, rot_h1t=sp.rot_h1t,
                      d_rot_h1t=torch.as_tensor(np.asarray(D) @ c, device="cuda").contiguous(), params=params)


def _green(tc: TangentCtx, w, dw):
    """(overlap, GT, dGT, dlog overlap) for restricted walkers (overlap = det(C^T w)^2), on the library's
    batched FP64 GEMM and inverse: m = C^T w, GT = w m^-1, dGT = (dw - GT dm) m^-1, dlog = 2 tr(m^-1 dm)."""
    m = _lib.gemm_batched(tc.C, w, opa="T")
    dm = _lib.gemm_batched(tc.C, dw, opa="T")
    minv, det = _lib.inverse_det(m)
    gt = _lib.gemm_batched(w, minv)
    x = dw.clone()
    _lib.gemm_batched(gt, dm, alpha=-1.0, beta=1.0, out=x)
    dgt = _lib.gemm_batched(x, minv)
    dlog = 2.0 * (minv * dm.transpose(1, 2)).sum((1, 2))
    return det * det, gt, dgt, dlog


def _energy(tc: TangentCtx, gt, dgt):
    """(E_L, dE_L) per walker: exact symmetric difference of the quadratic E_L(GT) plus the explicit term."""
    e = _lib.local_energy(tc.rot_chol, gt, tc.rot_h1t, tc.h0)
    big = float(gt.abs().max())
    small = float(dgt.abs().max())
    s = big / small if small > 0 else 1.0
    ep = _lib.local_energy(tc.rot_chol, (gt + s * dgt).contiguous(), tc.rot_h1t, tc.h0)
    em = _lib.local_energy(tc.rot_chol, (gt - s * dgt).contiguous(), tc.rot_h1t, tc.h0)
    de = (ep - em) / (2.0 * s) + 2.0 * (gt * tc.d_rot_h1t.to(C128)).sum((1, 2))
    return e, de


def init_state(tc: TangentCtx, w0: torch.Tensor) -> TangentState:
    """trot/prop/afqmc.py:19-76: weights 1, E_est = Re E_L(walker 0) (its derivative: the explicit term)."""
    W = w0.shape[0]
    dw0 = torch.zeros_like(w0)
    ov, gt, dgt, dlog = _green(tc, w0, dw0)
    e, de = _energy(tc, gt[:1].contiguous(), dgt[:1].contiguous())
    e0, de0 = float(e.real[0]), float(de.real[0])
    dev = w0.device
    return TangentState(w=w0.clone(), dw=dw0, weights=torch.ones(W, dtype=torch.float64, device=dev),
                        dweights=torch.zeros(W, dtype=torch.float64, device=dev), overlaps=ov, dlogov=dlog,
                        scal=torch.tensor([e0, e0], dtype=torch.float64, device=dev),
                        dscal=torch.tensor([de0, de0], dtype=torch.float64, device=dev),
                        nodes=torch.zeros(1, dtype=torch.int64, device=dev))


def tangent_step(tc: TangentCtx, st: TangentState, fields: torch.Tensor) -> TangentState:
    """One phaseless step (trot/prop/afqmc.py:82-144) with its lambda-derivative."""
    W, N, n = st.w.shape
    _, gt, dgt, _ = _green(tc, st.w, st.dw)
    fb = _lib.force_bias(gt, tc.rot_chol_t, 2.0)
    dfb = _lib.force_bias(dgt.contiguous(), tc.rot_chol_t, 2.0)
    xbar, shift, fbt = _lib.field_shift(fields, fb, tc.mf, tc.sqrt_dt)
    dfs = (-tc.sqrt_dt * 1j) * dfb
    dxbar = -dfs
    dshift = (dxbar * tc.mf[None, :]).sum(1)
    dfbt = (xbar * dfs).sum(1)
    v = _lib.vhs_build(xbar, tc.chol_flat).view(W, N, N)
    dv = _lib.vhs_build(dxbar.contiguous(), tc.chol_flat).view(W, N, N)
    vaug = torch.zeros((W, 2 * N, 2 * N), dtype=C128, device=st.w.device)
    vaug[:, :N, :N] = v
    vaug[:, :N, N:] = dv
    vaug[:, N:, N:] = v
    del v, dv
    waug = torch.cat([st.dw, st.w], dim=1).contiguous()
    waug = _lib.one_body_apply(tc.E_aug, waug)
    waug = _lib.taylor_apply(vaug, waug, tc.sqrt_dt, tc.n_terms, inplace=True)
    waug = _lib.one_body_apply(tc.E_aug, waug)
    del vaug
    dw_new, w_new = waug[:, :N].contiguous(), waug[:, N:].contiguous()
    ov_new, _, _, dlog_new = _green(tc, w_new, dw_new)

    # primal update on the HIP kernel (as the product's step), derivative alongside
    prm = tc.params
    weights = st.weights.clone()
    scal = st.scal.clone()
    nodes = st.nodes.clone()
    _lib.phaseless_update(weights, st.overlaps, ov_new, shift, fbt, scal, nodes, tc.dt, tc.h0_prop,
                          prm.weight_floor, prm.weight_cap, prm.pop_control_damping)
    with np.errstate(all="ignore"):
        ratio = ov_new / st.overlaps
        expo = -tc.sqrt_dt * shift + fbt + tc.dt * (st.scal[0] + tc.h0_prop)
        dexpo = -tc.sqrt_dt * dshift + dfbt + tc.dt * st.dscal[0]
        dlogratio = dlog_new - st.dlogov
        imp = torch.exp(expo) * ratio
        theta = torch.angle(torch.exp(-tc.sqrt_dt * shift) * ratio)
        aimp = imp.abs()
        p = aimp * torch.cos(theta)
        dtheta = (-tc.sqrt_dt * dshift + dlogratio).imag
        dp = aimp * ((dexpo + dlogratio).real * torch.cos(theta) - torch.sin(theta) * dtheta)
        keep = torch.isfinite(p) & (p >= prm.weight_floor) & (p <= prm.weight_cap)
        p = torch.where(keep, p, torch.zeros_like(p))
        dp = torch.where(keep, dp, torch.zeros_like(dp))
        wn = st.weights * p
        dwn = st.dweights * p + st.weights * dp
        dwn = torch.where(wn <= prm.weight_cap, dwn, torch.zeros_like(dwn))
        mean_w = float(weights.mean())
        dpop = -prm.pop_control_damping / tc.dt * float(dwn.mean()) / mean_w if mean_w > 1e-300 else 0.0
    dscal = torch.stack([torch.tensor(dpop, dtype=torch.float64, device=st.w.device), st.dscal[1]])
    return TangentState(w=w_new, dw=dw_new, weights=weights, dweights=dwn, overlaps=ov_new, dlogov=dlog_new,
                        scal=scal, dscal=dscal, nodes=nodes, step=st.step + 1)


def block_energy(tc: TangentCtx, st: TangentState):
    """trot/prop/blocks.py:236-257 (clamp at sqrt(2/dt) around E_est, weighted mean) with its derivative.
    The QR that precedes it (blocks.py:230) leaves GT, hence E_L, unchanged."""
    _, gt, dgt, _ = _green(tc, st.w, st.dw)
    e, de = _energy(tc, gt, dgt)
    e, de = e.real, de.real
    e_est, de_est = st.scal[1], st.dscal[1]
    clamp = ~torch.isfinite(e) | ((e - e_est).abs() > np.sqrt(2.0 / tc.dt))
    w = torch.where(torch.isfinite(e), st.weights, torch.zeros_like(st.weights))
    dw = torch.where(torch.isfinite(e), st.dweights, torch.zeros_like(st.dweights))
    es = torch.where(clamp, e_est.expand_as(e), e)
    des = torch.where(clamp, de_est.expand_as(de), de)
    sw = w.sum()
    eb = (w * es).sum() / sw
    deb = ((dw * es + w * des).sum() - eb * dw.sum()) / sw
    return float(eb), float(deb)


@dataclass
class Result:
    energy: float
    denergy: float
    checkpoints: List[TangentState] = field(default_factory=list)


def run(tc: TangentCtx, w0: torch.Tensor, fields: Callable[[int], torch.Tensor], n_steps: int,
        checkpoint_every: int = 0, resume: TangentState | None = None) -> Result:
    """n_steps tangent steps from w0 (or from a checkpoint), then the block energy and its derivative."""
    st = resume.clone() if resume is not None else init_state(tc, w0)
    cps = []
    while st.step < n_steps:
        st = tangent_step(tc, st, fields(st.step))
        if checkpoint_every and st.step % checkpoint_every == 0:
            cps.append(st.clone())
    e, de = block_energy(tc, st)
    return Result(e, de, cps)


class BlockEnergy(torch.autograd.Function):
    """E_block(lambda) with d/dlambda from the tangent propagation: ``E = BlockEnergy.apply(lam, fn)``,
    ``fn(lam_float) -> Result``; ``E.backward()`` leaves dE/dlambda in ``lam.grad``."""

    @staticmethod
    def forward(ctx, lam: torch.Tensor, fn):
        res = fn(float(lam))
        ctx.save_for_backward(torch.tensor(res.denergy, dtype=torch.float64))
        return torch.tensor(res.energy, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        (d,) = ctx.saved_tensors
        return grad * d, None
