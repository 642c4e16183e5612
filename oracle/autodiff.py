"""Differentiable CPU restatement of the oracle's phaseless block energy (torch, complex128 / float64).

TEST INFRASTRUCTURE ONLY.  Like the rest of ``oracle/``, only ``tests/`` may import this module; the product
path (``trot_amd``) never does.

The reference gets the response density dE_block/dh1 from ``jax.grad`` through ``afqmc_step`` and the block
energy (README.md:11; trot/prop/afqmc.py:79-144, trot/prop/blocks.py:236-257).  JAX is absent here, so this
module restates the same forward pass, function by function after ``oracle/afqmc_oracle.py`` (restricted
walkers, RHF trial, fields given), in torch operations that autograd can differentiate:

  build_prop_ctx   trot/prop/chol_afqmc_ops.py:74-80, 90-127     mf, h0_prop, h1_eff, E = expm(-dt/2 h1_eff)
  build_meas_ctx   trot/meas/rhf.py:308-319                      rot_h1 = C^H h1, rot_chol = C^H L_g
  half_green       trot/meas/rhf.py:21-25                        solve(m^T, w^T), m = C^H w
  force_bias       trot/meas/rhf.py:119-125                      2 rot_chol_flat vec(G)
  energy           trot/meas/rhf.py:41-60, 128-138               h0 + 2 G.rot_h1 + 2 sum c_g^2 - exchange
  shifts           trot/prop/afqmc.py:99-103                     fs, xbar, shift_term, fb_term
  trotter          trot/prop/chol_afqmc_ops.py:130-162, 198-208  E . taylor(i sqrt(dt) V, E w, K)
  overlap          trot/trial/rhf.py:45-47                       det(C^H w)^2
  update           trot/prop/afqmc.py:108-134                    |I| cos(theta), masks, weights, E_shift
  init             trot/prop/afqmc.py:19-76                      weights 1, E_est = E_shift = Re E_L(walker 0)
  block energy     trot/prop/blocks.py:236-257                   clamp at sqrt(2/dt) around E_est, weighted mean

The masks (importance floor and cap, weight cap, energy clamp) select with ``torch.where`` and carry no
derivative, as in the reference under jax.grad.  ``tests/test_autodiff_pins.py`` pins the forward value
against ``afqmc_oracle.block`` (the numpy oracle) and the gradient against central differences of that
oracle, so the autograd gradient stands in for the reference's jax.grad.
"""

from __future__ import annotations

import math

import numpy as np
import torch

C128 = torch.complex128
F64 = torch.float64


def _c(x) -> torch.Tensor:
    return torch.as_tensor(np.asarray(x)).to(C128)


def _half_green(w: torch.Tensor, cc: torch.Tensor) -> torch.Tensor:
    m = torch.einsum("ip,wiq->wpq", cc.conj(), w)
    return torch.linalg.solve(m.transpose(1, 2), w.transpose(1, 2))  # (W, n, N)


def _overlap(w: torch.Tensor, cc: torch.Tensor) -> torch.Tensor:
    m = torch.einsum("ip,wiq->wpq", cc.conj(), w)
    return torch.linalg.det(m) ** 2


def block_energy(h0: float, h1: torch.Tensor, chol, c, w0, fields, dt: float, n_terms: int = 6,
                 weight_floor: float = 1.0e-3, weight_cap: float = 100.0, damping: float = 0.1):
    """Block energy after ``len(fields)`` phaseless steps from walkers ``w0`` (W, N, n), as a torch scalar
    differentiable in ``h1`` (N, N float64).  Returns (energy, info) with info = counts of masked importances,
    capped weights and clamped energies (a comparison of derivatives is meaningful only where they are 0)."""
    chol_t = torch.as_tensor(np.asarray(chol), dtype=F64)
    G, N, _ = chol_t.shape
    cc = _c(c)
    cr = torch.as_tensor(np.asarray(c), dtype=F64)
    sqrt_dt = math.sqrt(dt)
    # build_prop_ctx: RHF total density 2 C C^H
    dm = 2.0 * (cr @ cr.T)
    mf = 1.0j * torch.einsum("gij,ji->g", chol_t, dm).to(C128)
    h0_prop = -h0 - 0.5 * torch.sum(mf * mf)
    v0m = 0.5 * torch.einsum("gik,gkj->ij", chol_t, chol_t)
    v1m = torch.einsum("g,gik->ik", (1.0j * mf).real, chol_t)
    he = h1 - v0m - v1m
    E = torch.linalg.matrix_exp(-0.5 * dt * he).to(C128)
    lflat = chol_t.reshape(G, N * N).to(C128)
    # build_meas_ctx (high memory mode)
    rot_h1 = (cr.T @ h1).to(C128)
    rot_chol = torch.einsum("ip,gij->gpj", cr, chol_t).to(C128)
    rot_chol_flat = rot_chol.reshape(G, -1)

    def energy(w):
        g = _half_green(w, cc)
        e1 = 2.0 * torch.einsum("wij,ij->w", g, rot_h1)
        f = torch.einsum("gij,wkj->wgik", rot_chol, g)
        cg = torch.einsum("wgii->wg", f)
        exc = torch.einsum("wgij,wgji->w", f, f)
        return h0 + e1 + 2.0 * torch.sum(cg * cg, dim=1) - exc

    w = _c(w0)
    W = w.shape[0]
    weights = torch.ones(W, dtype=F64)
    ov = _overlap(w, cc)
    e_est = energy(w[:1]).real.mean()
    pop = e_est
    info = {"masked_importance": 0, "capped_weight": 0, "clamped_energy": 0}
    for x in fields:
        x = torch.as_tensor(np.asarray(x), dtype=F64).to(C128)
        g = _half_green(w, cc)
        fb = 2.0 * g.reshape(W, -1) @ rot_chol_flat.T
        fs = -sqrt_dt * (1.0j * fb - mf)
        xbar = x - fs
        shift_term = torch.sum(xbar * mf, dim=1)
        fb_term = torch.sum(x * fs - 0.5 * fs * fs, dim=1)
        v = (xbar @ lflat).reshape(W, N, N)
        u = E @ w
        term, acc = u, u
        for k in range(1, n_terms + 1):
            term = (1.0j * sqrt_dt / k) * (v @ term)
            acc = acc + term
        w = E @ acc
        ov_new = _overlap(w, cc)
        ratio = ov_new / ov
        exponent = -sqrt_dt * shift_term + fb_term + dt * (pop + h0_prop)
        imp = torch.exp(exponent) * ratio
        theta = torch.angle(torch.exp(-sqrt_dt * shift_term) * ratio)
        p = torch.abs(imp) * torch.cos(theta)
        bad = ~torch.isfinite(p) | (p < weight_floor) | (p > weight_cap)
        info["masked_importance"] += int(bad.sum())
        p = torch.where(bad, torch.zeros_like(p), p)
        wn = weights * p
        cap = wn > weight_cap
        info["capped_weight"] += int(cap.sum())
        weights = torch.where(cap, torch.zeros_like(wn), wn)
        pop = e_est - damping * torch.log(torch.clamp(weights.mean(), min=1.0e-300)) / dt
        ov = ov_new
    e = energy(w).real
    finite = torch.isfinite(e)
    clamp = ~finite | ((e - e_est).abs() > math.sqrt(2.0 / dt))
    info["clamped_energy"] = int(clamp.sum())
    wts = torch.where(finite, weights, torch.zeros_like(weights))
    es = torch.where(clamp, e_est.expand_as(e), e)
    return torch.sum(wts * es) / torch.sum(wts), info
