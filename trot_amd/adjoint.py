"""Reverse-mode (adjoint) derivative of the block energy with respect to the one-body Hamiltonian h1.

The reference is differentiable end to end (README.md:11): jax.grad of the block energy after K phaseless
steps (trot/prop/afqmc.py:79-144, trot/prop/blocks.py:236-257) with respect to h1 gives the response density
dE/dh1, and dE/dlambda = <dE/dh1, D> for any one-body perturbation h1 + lambda D.  This module is that
gradient written by hand for restricted walkers and an RHF trial, on the library's own kernels
(trot_gemm_batched, trot_inverse_det, the force-bias and VHS GEMMs, the one-body kernel, the local energy):

forward (one step, as trot/prop/afqmc.py:82-144)
  m = C^T w, GT = w m^-1;  fb = 2 R' GT;  xbar, shift_term, fb_term (afqmc.py:99-103);  V = sum_g xbar_g L_g;
  u = E w;  T_l = (i sqrt(dt) V)^l u / l!;  t = sum_l T_l;  w' = E t;  ratio = det(C^T w')^2 / ov;
  p = exp(Re fb_term + dt Re(E_shift + h0_prop)) Re(exp(-sqrt(dt) shift_term) ratio) = |I| cos(theta),
  masked as afqmc.py:122-130;  weight' = weight p.
reverse (cotangents x^ with dL = Re(conj(x^) dx) for complex x)
  p^ = weight^' weight, weight^ = weight^' p (kept walkers);  ratio^, shift_term^, fb_term^ from p;
  w'^ += 2 l'^ C m'^-H (the new overlap),  E^ += w'^ t^H,  t^ = E w'^;
  u^ = sum_j y_j / j!, y_j = (conj(a) V^H)^j t^ (V symmetric: V^H = conj V);
  conj(V^) = a sum_j conj(y_j) z_j^T, z_j = sum_l T_l l! / (j + l + 1)!  ->  xbar^_g = <L_g, V^>;
  E^ += u^ w^H,  w^ = E u^;  fb^ = i sqrt(dt) fs^ with fs^ = -xbar^ + conj(x - fs) fb_term^;
  GT^ = 2 R'^T fb^;  w^ += [GT^ - C (GT^H GT^ - 2 l^ I)] m^-H with l^ the old overlap's cotangent.
  h1^ = -dt/2 L*(A, Re E^), the adjoint Frechet derivative of E = exp(A), A = -dt/2 h1_eff (the top-right
  block of exp([[A^T, Re E^], [0, A^T]])), plus the explicit one-body energy term 2 Re(sum_w e^_w GT_w C^T)^T.
The block energy's E_L (trot/meas/rhf.py:128-138) gives GT^ = conj(dE_L/dGT) e^ with
  dE_L/dGT = 2 rot_h1^T + 4 sum_g c_g R'_g - 2 sum_g rot_chol_g^T f_g^T,  f_g = rot_chol_g GT, c_g = tr f_g.
The population-control shift scales every weight alike and cancels in the block energy; masked importances
and capped weights carry no derivative (as the tangent propagation in trot_amd.diff).

Memory: the forward pass keeps the state every ``checkpoint_every`` steps; the reverse pass recomputes each
segment from its checkpoint (one more forward pass) and walks it backwards.  ``precision="f32"`` runs the
Cholesky contractions (force bias, VHS and their adjoints, the dominant flops) on FP32 MFMA with FP32
accumulation, complex128 walkers; every other product is FP64.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, List

import numpy as np
import scipy.linalg
import torch

from . import _lib, meas, prop, trial as trial_mod
from .core import HamChol, System

C128 = torch.complex128
F64 = torch.float64


@dataclass
class AdjCtx:
    """What the forward and reverse steps need at one h1."""

    N: int
    n: int
    G: int
    dt: float
    sqrt_dt: float
    n_terms: int
    h0: float
    h0_prop: complex
    mf: torch.Tensor  # (G,) complex128
    E: torch.Tensor  # (N, N) float64, exp(-dt/2 h1_eff)
    A: np.ndarray  # -dt/2 h1_eff (host)
    C: torch.Tensor  # (N, n) float64 trial
    Lsym: torch.Tensor  # (G, symblk_cols(N)) float64: the stored 8 x 8 blocks of every symmetric L_g
    sym_inv: torch.Tensor  # (N*N,) long: stored column of entry (i, k) of a symmetric matrix
    sym_i1: torch.Tensor  # (cols,) long: entry (i, k) a stored column holds
    sym_i2: torch.Tensor  # (cols,) long: its mirror (k, i)
    sym_off: torch.Tensor  # (cols,) float64: 1 for columns of off-diagonal blocks, 0 on diagonal blocks
    sym_valid: torch.Tensor  # (cols,) float64: 0 for the padding of the last block row
    Rt: torch.Tensor  # (G, N*n) float64: R'[g, j*n + p] = rot_chol[g, p, j]
    rot_chol: torch.Tensor  # (G, n, N) float64
    rot_h1t: torch.Tensor  # (N, n) float64 = h1 C
    params: prop.QmcParams
    precision: str = "f64"
    Rt32: torch.Tensor | None = None
    Lsym32: torch.Tensor | None = None


def _symblk_maps(N: int, device):
    """Index maps between a dense symmetric N x N matrix and its sym-block storage (trot_pack_symblk's
    layout): packing the matrices of flat indices (+1, so that padding reads 0) and of their transposes."""
    idx = torch.arange(1, N * N + 1, dtype=F64, device=device).view(1, N, N)
    p1 = _lib.pack_symblk(idx)[0]
    p2 = _lib.pack_symblk(idx[0].T.contiguous().view(1, N, N))[0]
    valid = p1 > 0
    i1 = (p1 - 1).clamp(min=0).long()
    i2 = (p2 - 1).clamp(min=0).long()
    off = valid & ((i1 // N) // 8 != (i1 % N) // 8)
    cols = torch.arange(p1.numel(), device=device)
    inv = torch.zeros(N * N, dtype=torch.long, device=device)
    inv[i1[valid]] = cols[valid]
    inv[i2[valid]] = cols[valid]  # V is symmetric: either stored copy of a diagonal-block entry serves
    return inv, i1, i2, off.to(F64), valid.to(F64)


def build_ctx(h0: float, h1, chol, c: np.ndarray, params: prop.QmcParams, precision: str = "f64") -> AdjCtx:
    """Contexts of the product's builders (trot_amd.prop / trot_amd.meas) at this h1."""
    if precision not in ("f64", "f32"):
        raise ValueError(f"precision must be 'f64' or 'f32', got {precision!r}")
    h1 = h1.detach().cpu().numpy() if isinstance(h1, torch.Tensor) else np.asarray(h1, dtype=np.float64)
    N, n = c.shape
    ham = HamChol(h0, h1, chol).to_device()
    sys_ = System(N, (n, n), "restricted")
    tr = trial_mod.RhfTrial(c)
    tops = trial_mod.make_rhf_trial_ops(sys_)
    pctx = prop.make_prop_ops("restricted", "restricted").build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = meas.make_rhf_meas_ops(sys_).build_meas_ctx(ham, tr)
    G = ham.chol.shape[0]
    # h1_eff = h1 - 1/2 sum_g L_g L_g - sum_g Re(i mf_g) L_g (trot/prop/chol_afqmc_ops.py:90-100), the generator
    # of E, from the same device contractions the propagation context uses
    _, mfl, v0m = _lib.build_prop_tensors(ham.chol, prop._total_dm(tops.get_rdm1(tr)))
    h1_eff = torch.as_tensor(h1, device="cuda") - v0m - (1j * mfl).real
    sp = mctx.spin
    ac = AdjCtx(N=N, n=n, G=G, dt=pctx.dt, sqrt_dt=pctx.sqrt_dt, n_terms=params.n_exp_terms, h0=float(h0),
                h0_prop=complex(pctx.h0_prop), mf=pctx.mf_shifts, E=pctx.exp_h1_half.contiguous(),
                A=(-0.5 * pctx.dt) * h1_eff.cpu().numpy(), C=torch.as_tensor(c, dtype=F64, device="cuda"),
                Lsym=_lib.pack_symblk(ham.chol), sym_inv=None, sym_i1=None, sym_i2=None,
                sym_off=None, sym_valid=None, Rt=sp.rot_chol_t.reshape(G, N * n), rot_chol=sp.rot_chol,
                rot_h1t=sp.rot_h1t, params=params, precision=precision)
    ac.sym_inv, ac.sym_i1, ac.sym_i2, ac.sym_off, ac.sym_valid = _symblk_maps(N, ac.Lsym.device)
    if precision == "f32":
        ac.Rt32 = ac.Rt.to(torch.float32).contiguous()
        ac.Lsym32 = ac.Lsym.to(torch.float32).contiguous()
    return ac


@dataclass
class Snap:
    """Primal state between steps."""

    w: torch.Tensor  # (W, N, n) complex128
    weights: torch.Tensor  # (W,)
    overlaps: torch.Tensor  # (W,) complex128
    pop: float
    e_est: float
    step: int = 0


# ------------------------------------------------------------------------------------------------ pieces


def _green(ac: AdjCtx, w: torch.Tensor):
    """(m^-1, det m, GT) with m = C^T w."""
    m = _lib.gemm_batched(ac.C, w, opa="T")
    minv, det = _lib.inverse_det(m)
    return minv, det, _lib.gemm_batched(w, minv)


def _fb(ac: AdjCtx, gt: torch.Tensor, alpha: float) -> torch.Tensor:
    """alpha R' GT (W, G)."""
    W = gt.shape[0]
    if ac.precision == "f32":
        return _lib.force_bias_lp(gt, ac.Rt32.view(ac.G, ac.N, ac.n), alpha, "f32")
    return _lib.force_bias_flat(gt.reshape(W, ac.N * ac.n), ac.Rt, alpha)


def _rt_apply(ac: AdjCtx, x: torch.Tensor) -> torch.Tensor:
    """sum_g x_g R'_g as (W, N, n): the adjoint of the force-bias contraction."""
    if ac.precision == "f32":
        y = _lib.vhs_build_lp(x.contiguous(), ac.Rt32, "f32", out_f64=True)
    else:
        y = _lib.vhs_build(x.contiguous(), ac.Rt)
    return y.view(x.shape[0], ac.N, ac.n)


def _vhs(ac: AdjCtx, xbar: torch.Tensor) -> torch.Tensor:
    """V = sum_g xbar_g L_g (W, N, N): the product over the stored half of the symmetric L_g (sym-block
    storage, half the flops of the dense product), then unpacked."""
    W = xbar.shape[0]
    if ac.precision == "f32":
        v = _lib.vhs_build_lp(xbar, ac.Lsym32, "f32", out_f64=True)
    else:
        v = _lib.vhs_build(xbar, ac.Lsym)
    return v.index_select(1, ac.sym_inv).view(W, ac.N, ac.N)


def _l_dot(ac: AdjCtx, vhat_conj: torch.Tensor) -> torch.Tensor:
    """xbar^_g = <L_g, V^> = conj(sum_ij L_g,ij conj(V^)_ij), (W, G).  L_g is symmetric, so the sum runs over
    its stored blocks: P = conj(V^)_ik + conj(V^)_ki on off-diagonal blocks and conj(V^)_ik on diagonal ones,
    contracted with the sym-block L (half the flops of the dense contraction)."""
    W = vhat_conj.shape[0]
    flat = vhat_conj.reshape(W, ac.N * ac.N)
    P = (flat.index_select(1, ac.sym_i1) + ac.sym_off * flat.index_select(1, ac.sym_i2)) * ac.sym_valid
    cols = P.shape[1]
    if ac.precision == "f32":
        r = _lib.force_bias_lp(P.view(W, cols // 64, 64), ac.Lsym32.view(ac.G, cols // 64, 64), 1.0, "f32")
    else:
        r = _lib.force_bias_flat(P, ac.Lsym, 1.0)
    return r.conj()


def _taylor_terms(ac: AdjCtx, V: torch.Tensor, u: torch.Tensor) -> List[torch.Tensor]:
    """T_l = (i sqrt(dt) V)^l u / l!, l = 0 .. K (trot/prop/utils.py:20-29: term <- (a / l) V term)."""
    terms = [u]
    for l in range(1, ac.n_terms + 1):
        terms.append(_lib.gemm_batched(V, terms[-1], alpha=1j * ac.sqrt_dt / l))
    return terms


def _importance(ac: AdjCtx, snap: Snap, shift, fbt, ov_new):
    """(p, keep, pre, z, ratio): p = pre Re(z) = |I| cos(theta) (trot/prop/afqmc.py:111-124)."""
    ratio = ov_new / snap.overlaps
    z = torch.exp(-ac.sqrt_dt * shift) * ratio
    pre = torch.exp(fbt.real + ac.dt * (snap.pop + ac.h0_prop.real))
    p = pre * z.real
    prm = ac.params
    keep = torch.isfinite(p) & (p >= prm.weight_floor) & (p <= prm.weight_cap)
    return torch.where(keep, p, torch.zeros_like(p)), keep, pre, z, ratio


def _step_parts(ac: AdjCtx, snap: Snap, x: torch.Tensor):
    minv, det, gt = _green(ac, snap.w)
    fb = _fb(ac, gt, 2.0)
    xbar, shift, fbt = _lib.field_shift(x, fb, ac.mf, ac.sqrt_dt)
    V = _vhs(ac, xbar)
    u = _lib.one_body_apply(ac.E, snap.w)
    terms = _taylor_terms(ac, V, u)
    t = terms[0].clone()
    for tl in terms[1:]:
        t += tl
    w1 = _lib.one_body_apply(ac.E, t)
    minv1, det1, _ = _green(ac, w1)
    return dict(minv=minv, gt=gt, fb=fb, xbar=xbar, shift=shift, fbt=fbt, V=V, u=u, terms=terms, t=t, w1=w1,
                minv1=minv1, ov1=det1 * det1)


# ------------------------------------------------------------------------------------------------ forward


def init_snap(ac: AdjCtx, w0: torch.Tensor) -> Snap:
    """trot/prop/afqmc.py:19-76: weights 1, overlaps, E_est = E_shift = Re E_L(walker 0)."""
    _, det, gt = _green(ac, w0)
    e0 = float(_lib.local_energy(ac.rot_chol, gt[:1].contiguous(), ac.rot_h1t, ac.h0).real[0])
    W = w0.shape[0]
    return Snap(w=w0.clone(), weights=torch.ones(W, dtype=F64, device=w0.device), overlaps=det * det, pop=e0,
                e_est=e0)


def step(ac: AdjCtx, snap: Snap, x: torch.Tensor) -> Snap:
    """One phaseless step (trot/prop/afqmc.py:82-144) on the adjoint's own kernels."""
    d = _step_parts(ac, snap, x)
    p, keep, *_ = _importance(ac, snap, d["shift"], d["fbt"], d["ov1"])
    prm = ac.params
    wn = snap.weights * p
    wn = torch.where(wn > prm.weight_cap, torch.zeros_like(wn), wn)
    mean_w = max(float(wn.mean()), 1.0e-300)
    pop = snap.e_est - prm.pop_control_damping * math.log(mean_w) / ac.dt
    return Snap(w=d["w1"], weights=wn, overlaps=d["ov1"], pop=pop, e_est=snap.e_est, step=snap.step + 1)


def _block_energy_terms(ac: AdjCtx, snap: Snap):
    """trot/prop/blocks.py:236-257: clamp at sqrt(2/dt) around E_est, weighted mean (QR leaves GT unchanged)."""
    minv, _, gt = _green(ac, snap.w)
    e = _lib.local_energy(ac.rot_chol, gt, ac.rot_h1t, ac.h0).real
    finite = torch.isfinite(e)
    clamp = ~finite | ((e - snap.e_est).abs() > math.sqrt(2.0 / ac.dt))
    wts = torch.where(finite, snap.weights, torch.zeros_like(snap.weights))
    es = torch.where(clamp, torch.full_like(e, snap.e_est), e)
    S = wts.sum()
    eb = float((wts * es).sum() / S)
    return eb, gt, minv, finite, clamp, wts, es, S


# ------------------------------------------------------------------------------------------------ reverse


def _green_adjoint(ac: AdjCtx, gt, minv, gt_hat, l_hat, w_hat):
    """w^ += [GT^ - C (GT^H GT^ - 2 l^ I)] m^-H for GT = w m^-1, m = C^T w, l = log det(m)^2."""
    X = _lib.gemm_batched(gt, gt_hat, opa="C")
    if l_hat is not None:
        X.diagonal(dim1=1, dim2=2).sub_(2.0 * l_hat[:, None])
    Y = gt_hat.clone()
    _lib.gemm_batched(ac.C, X, alpha=-1.0, beta=1.0, out=Y)
    return _lib.gemm_batched(Y, minv, opb="C", beta=1.0, out=w_hat)


def _k_slices(W: int, tiles: int, target_wg: int = 512) -> int:
    """Walker slices for a split-K outer product: the largest divisor of W up to target_wg / tiles."""
    s = max(1, min(W, target_wg // max(tiles, 1)))
    while W % s:
        s -= 1
    return s


def _acc_outer(acc: torch.Tensor, x: torch.Tensor, y: torch.Tensor) -> None:
    """acc (N x N complex) += sum_w x_w y_w^H, a GEMM over K = W n (walkers moved to the column index).  The N x N
    output has few tiles, so K is split over S walker slices (one batched GEMM of S partial products, about 512
    workgroups) and the partial products are summed over the slices."""
    W, N, n = x.shape
    xp = x.permute(1, 0, 2).reshape(N, W * n)
    yp = y.permute(1, 0, 2).reshape(N, W * n)
    S = _k_slices(W, ((N + 63) // 64) ** 2)
    if S == 1:
        _lib.gemm_batched(xp, yp, opb="C", beta=1.0, out=acc.view(1, N, N))
        return
    kc = (W // S) * n
    part = _lib.gemm_batched(xp.view(N, S, kc).permute(1, 0, 2), yp.view(N, S, kc).permute(1, 0, 2), opb="C")
    acc += part.sum(0)


def _energy_grad_gt(ac: AdjCtx, gt: torch.Tensor, chunk: int = 64) -> torch.Tensor:
    """dE_L/dGT (holomorphic) of trot/meas/rhf.py:41-60, 128-138 for every walker, (W, N, n)."""
    W, N, n = gt.shape
    G = ac.G
    c = _fb(ac, gt, 1.0)  # c_g = tr(rot_chol_g GT)
    out = 4.0 * _rt_apply(ac, c)
    out += 2.0 * ac.rot_h1t.to(C128)[None]
    rc_all = ac.rot_chol.reshape(G * n, N)
    for w0 in range(0, W, chunk):
        g = gt[w0:w0 + chunk]
        f = _lib.gemm_batched(rc_all, g)  # f[(g, p), q] = (rot_chol_g GT)[p, q]
        h = f.view(-1, G, n, n).transpose(-1, -2).contiguous().view(-1, G * n, n)
        _lib.gemm_batched(rc_all, h, opa="T", alpha=-2.0, beta=1.0, out=out[w0:w0 + chunk])
    return out


def _step_adjoint(ac: AdjCtx, snap: Snap, x: torch.Tensor, w_hat: torch.Tensor, wt_hat: torch.Tensor,
                  e_hat: torch.Tensor):
    """Cotangents of (w, weights) before the step from those after it; accumulates E^ into e_hat."""
    d = _step_parts(ac, snap, x)
    prm = ac.params
    p, keep, pre, z, ratio = _importance(ac, snap, d["shift"], d["fbt"], d["ov1"])
    wn = snap.weights * p
    live = keep & (wn <= prm.weight_cap)
    p_hat = torch.where(live, wt_hat * snap.weights, torch.zeros_like(wt_hat))
    wt_prev = torch.where(live, wt_hat * p, torch.zeros_like(wt_hat))
    fbt_hat = (p_hat * p).to(C128)
    z_hat = (p_hat * pre).to(C128)
    shift_hat = torch.conj(-ac.sqrt_dt * z) * z_hat
    ratio_hat = torch.conj(torch.exp(-ac.sqrt_dt * d["shift"])) * z_hat
    l_new = torch.conj(ratio) * ratio_hat
    l_new = torch.where(live, l_new, torch.zeros_like(l_new))  # a dead walker's ratio may not be finite
    shift_hat = torch.where(live, shift_hat, torch.zeros_like(shift_hat))
    # the new overlap: w'^ += 2 l'^ C m'^-H
    cm = _lib.gemm_batched(ac.C, d["minv1"], opb="C")
    w_hat = w_hat + 2.0 * l_new[:, None, None] * cm
    # w' = E t
    _acc_outer(e_hat, w_hat, d["t"])
    t_hat = _lib.one_body_apply(ac.E, w_hat)
    # t = sum_l T_l: y_j = (conj(a) V^H)^j t^, u^ = sum_j y_j / j!
    K = ac.n_terms
    W, N, n = snap.w.shape
    a = 1j * ac.sqrt_dt
    Y = torch.empty((W, N, K * n), dtype=C128, device=w_hat.device)
    Z = torch.empty((W, N, K * n), dtype=C128, device=w_hat.device)
    Y[:, :, :n] = t_hat
    u_hat = t_hat.clone()
    y = t_hat
    for j in range(1, K + 1):
        y = _lib.gemm_batched(d["V"], y, opa="C", alpha=np.conj(a))
        u_hat += y / math.factorial(j)
        if j < K:
            Y[:, :, j * n:(j + 1) * n] = y
    terms = d["terms"]
    for j in range(K):
        zj = terms[0] * (1.0 / math.factorial(j + 1))
        for l in range(1, K - j):
            zj = zj + terms[l] * (math.factorial(l) / math.factorial(j + l + 1))
        Z[:, :, j * n:(j + 1) * n] = zj
    vhat_c = _lib.gemm_batched(Y, Z, opa="R", opb="T", alpha=a)  # conj(V^)
    xbar_hat = _l_dot(ac, vhat_c)
    del Y, Z, vhat_c
    # u = E w
    _acc_outer(e_hat, u_hat, snap.w)
    w_prev = _lib.one_body_apply(ac.E, u_hat)
    # shifted fields (trot/prop/afqmc.py:99-103)
    fs = x.to(C128) - d["xbar"]
    xbar_hat = xbar_hat + torch.conj(ac.mf)[None, :] * shift_hat[:, None]
    fs_hat = -xbar_hat + torch.conj(x.to(C128) - fs) * fbt_hat[:, None]
    fb_hat = (1j * ac.sqrt_dt) * fs_hat
    gt_hat = 2.0 * _rt_apply(ac, fb_hat)
    l_old = torch.where(live, -l_new, torch.zeros_like(l_new))
    w_prev = _green_adjoint(ac, d["gt"], d["minv"], gt_hat, l_old, w_prev)
    return w_prev, wt_prev


def _frechet_adjoint(A: np.ndarray, Ebar: np.ndarray) -> np.ndarray:
    """L*(A, Ebar) = L(A^T, Ebar): the top-right block of exp([[A^T, Ebar], [0, A^T]])."""
    N = A.shape[0]
    big = scipy.linalg.expm(np.block([[A.T, Ebar], [np.zeros((N, N)), A.T]]))
    return big[:N, N:]


@dataclass
class Result:
    energy: float
    grad_h1: torch.Tensor  # (N, N) float64 symmetric: dE_block / dh1
    n_steps: int
    checkpoints: int


def run(ac: AdjCtx, w0: torch.Tensor, fields: Callable[[int], torch.Tensor], n_steps: int,
        checkpoint_every: int = 20) -> Result:
    """Block energy after n_steps steps from w0 and its gradient with respect to h1 (restricted walkers)."""
    k = max(1, int(checkpoint_every))
    snap = init_snap(ac, w0)
    e0_snap = snap
    cps = {0: snap}
    for s in range(n_steps):
        snap = step(ac, snap, fields(s))
        if (s + 1) % k == 0 and s + 1 < n_steps:
            cps[s + 1] = snap
    eb, gt, minv, finite, clamp, wts, es, S = _block_energy_terms(ac, snap)
    N, n = ac.N, ac.n
    wt_hat = torch.where(finite, (es - eb) / S, torch.zeros_like(es))
    e_hat = torch.where(clamp, torch.zeros_like(wts), wts / S)
    # explicit one-body term of E_L: 2 Re(sum_w e^_w GT_w C^T)^T, and of E_est through clamped walkers
    gbar = (e_hat.to(C128)[:, None, None] * gt).sum(0)
    est_hat = float((torch.where(clamp & finite, wts, torch.zeros_like(wts))).sum() / S)
    if est_hat != 0.0:
        _, _, gt0 = _green(ac, e0_snap.w[:1].contiguous())
        gbar = gbar + est_hat * gt0[0]
    h1_hat = 2.0 * _lib.gemm_batched(ac.C, gbar, opb="T").real[0]
    gt_hat = torch.conj(_energy_grad_gt(ac, gt)) * e_hat.to(C128)[:, None, None]
    w_hat = _green_adjoint(ac, gt, minv, gt_hat, None, torch.zeros_like(gt))
    e_acc = torch.zeros((N, N), dtype=C128, device=w0.device)
    for s0 in sorted(cps, reverse=True):
        s1 = min(s0 + k, n_steps)
        seg = [cps[s0]]
        for s in range(s0, s1 - 1):
            seg.append(step(ac, seg[-1], fields(s)))
        for s in range(s1 - 1, s0 - 1, -1):
            w_hat, wt_hat = _step_adjoint(ac, seg[s - s0], fields(s), w_hat, wt_hat, e_acc)
        del seg
    h1_hat = h1_hat.cpu().numpy() + (-0.5 * ac.dt) * _frechet_adjoint(ac.A, e_acc.real.cpu().numpy())
    g = 0.5 * (h1_hat + h1_hat.T)
    return Result(eb, torch.as_tensor(g), n_steps, len(cps))


class ResponseEnergy(torch.autograd.Function):
    """E_block(h1) with the adjoint as its backward: ``E = ResponseEnergy.apply(h1, fn)`` with
    ``fn(h1 (N, N) float64 tensor) -> Result``; ``E.backward()`` leaves dE/dh1 (the response density) in
    ``h1.grad``, and dE/dlambda in ``lam.grad`` when h1 = h1_0 + lam D was built from a leaf ``lam``."""

    @staticmethod
    def forward(ctx, h1: torch.Tensor, fn):
        res = fn(h1.detach())
        ctx.save_for_backward(res.grad_h1.to(dtype=h1.dtype, device=h1.device))
        return torch.tensor(res.energy, dtype=h1.dtype, device=h1.device)

    @staticmethod
    def backward(ctx, grad):
        (g,) = ctx.saved_tensors
        return grad * g, None
