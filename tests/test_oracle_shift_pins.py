"""Pins of the oracle's field-shift and importance-weight algebra (trot/prop/afqmc.py:99-134).

The oracle restates fs = -sqrt(dt)(i fb - mf), xbar = x - fs, shift_term = sum xbar mf,
fb_term = sum (x fs - fs^2/2), exponent = -sqrt(dt) shift_term + fb_term + dt (E_shift + h0_prop) and
theta = angle(exp(-sqrt(dt) shift_term) ratio).  Nothing here retypes those formulas; each check is a fact
the mathematics fixes:

* importance sampling is unbiased: for ANY complex force bias, the weighted propagated walker
  E_x[exp(exponent) B(xbar) w] is the same matrix (Gauss-Hermite quadrature, exact to rounding);
* the weighted propagator is the imaginary-time propagator: with a nonzero mean field and force bias,
  E_x[exp(exponent) |B(xbar) w>] = exp(-dt (H - E_shift)) |w> in Fock space up to O(dt^2)
  (H from oracle/fock.py, never from the oracle's h1_eff / h0_prop);
* the force-bias shift is the optimal one: d ln I / dx at x = 0 vanishes to O(dt^1.5) (a wrong shift
  leaves an O(sqrt(dt)) slope);
* the mean-field shift is the trial's own force bias: mf_g = i <T|L_g|T> / <T|T> (Fock space);
* theta known answers worked by hand with shift_term != 0, where theta of exp(-sqrt(dt) shift_term) ratio
  and theta of the full exponent differ.

Each plausible slip (sign of xbar or fs, a lost 1/2, shift_term from x, theta from the full exponent,
h0_prop or h1_eff sign, mf sign) is planted in the oracle here on the CPU and must fail at least one check.
"""

from __future__ import annotations

import math
from dataclasses import replace

import numpy as np
import pytest
import scipy.linalg

from oracle import afqmc_oracle as O
from oracle import fock as F

N, NOCC, G = 4, 2, 2
H0 = 0.3
E_SHIFT = 0.7


def _system():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((N, N))
    h1 = 0.3 * (a + a.T)
    b = rng.standard_normal((G, N, N))
    chol = 0.4 * (b + np.swapaxes(b, 1, 2))
    c = np.linalg.qr(rng.standard_normal((N, NOCC)))[0]
    w = np.linalg.qr(rng.standard_normal((N, NOCC)) + 1j * rng.standard_normal((N, NOCC)))[0]
    fbs = [rng.standard_normal(G) + 1j * rng.standard_normal(G), 3.0 * (rng.standard_normal(G)
                                                                          + 1j * rng.standard_normal(G))]
    return h1, chol, c, w, fbs


def _quadrature(nq: int):
    """Tensor Gauss-Hermite rule for E over G independent N(0,1) variables."""
    x, wt = np.polynomial.hermite_e.hermegauss(nq)
    wt = wt / math.sqrt(2.0 * math.pi)
    xs = np.meshgrid(*([x] * G), indexing="ij")
    ws = np.meshgrid(*([wt] * G), indexing="ij")
    return np.stack([g.ravel() for g in xs], 1), np.prod(np.stack([g.ravel() for g in ws], 1), 1)


def _weighted(ctx, w, fb, nq):
    """Quadrature nodes' weights, importance exponents and propagated walkers for force bias fb."""
    X, wq = _quadrature(nq)
    q = X.shape[0]
    xbar, st, ft = O.shift_terms(X, np.repeat(fb[None], q, 0), ctx)
    ex = O.importance_exponent(st, ft, E_SHIFT, ctx)
    wp = O.apply_trotter_r(np.repeat(w[None], q, 0), xbar, ctx, 6)
    return wq, ex, wp


# ----------------------------------------------------------------------------------------------- checks


def check_importance_sampling_is_unbiased() -> bool:
    h1, chol, c, w, fbs = _system()
    ctx = O.build_prop_ctx(H0, h1, chol, O.rhf_rdm1(c), 0.02)
    mats = []
    for fb in [-1j * ctx.mf_shifts] + fbs:  # the first gives fs = 0 for the correct shift
        wq, ex, wp = _weighted(ctx, w, fb, 40)
        mats.append(np.einsum("q,q,qij->ij", wq, np.exp(ex), wp))
    ref = mats[0]
    return all(np.max(np.abs(m - ref)) <= 1e-12 * np.max(np.abs(ref)) for m in mats[1:])


def _fock_error(dt: float) -> float:
    h1, chol, c, w, fbs = _system()
    fsp = F.FockSpace(N, (NOCC, NOCC))
    ham = fsp.hamiltonian(H0, h1, chol)
    ctx = O.build_prop_ctx(H0, h1, chol, O.rhf_rdm1(c), dt)
    wq, ex, wp = _weighted(ctx, w, fbs[0], 30)
    est = sum(wq[k] * np.exp(ex[k]) * fsp.state(wp[k], wp[k]) for k in range(wq.size))
    tgt = scipy.linalg.expm(-dt * (ham - E_SHIFT * np.eye(ham.shape[0]))) @ fsp.state(w, w)
    return float(np.linalg.norm(est - tgt) / np.linalg.norm(tgt))


def check_weighted_propagator_is_exp_minus_dt_h() -> bool:
    e1, e2 = _fock_error(0.02), _fock_error(0.01)
    # measured with the oracle as written: 4.4e-4 and 1.1e-4 (ratio 3.9, O(dt^2)); O(dt) slips give ratio ~2
    return e1 < 1e-3 and e2 < e1 / 3.2


def _log_importance_slope(dt: float):
    h1, chol, c, w, _ = _system()
    ctx = O.build_prop_ctx(H0, h1, chol, O.rhf_rdm1(c), dt)
    fb = O.force_bias_rw(w[None], O.build_meas_ctx_rhf(h1, chol, c), c)[0]
    ov0 = O.overlap_r(w[None], c)[0]
    eps = 1e-5
    slope = []
    for k in range(G):
        vals = []
        for s in (1.0, -1.0):
            x = np.zeros((1, G))
            x[0, k] = s * eps
            xbar, st, ft = O.shift_terms(x, fb[None], ctx)
            ex = O.importance_exponent(st, ft, 0.0, ctx)
            wp = O.apply_trotter_r(w[None], xbar, ctx, 6)
            vals.append(np.log(np.exp(ex[0]) * O.overlap_r(wp, c)[0] / ov0))
        slope.append((vals[0] - vals[1]) / (2 * eps))
    return np.abs(np.array(slope)), math.sqrt(dt) * np.abs(ctx.mf_shifts - 1j * fb)


def check_force_bias_shift_is_optimal() -> bool:
    s1, scale = _log_importance_slope(0.01)
    s2, _ = _log_importance_slope(0.0025)
    # the slope cancels to O(dt^1.5): quartering dt divides it by ~8 (an O(sqrt(dt)) slip: by ~2)
    return bool(np.all(s1 < 0.1 * scale) and np.all(s2 < s1 / 5.0))


def check_mean_field_is_trial_force_bias() -> bool:
    h1, chol, c, _, _ = _system()
    ctx = O.build_prop_ctx(H0, h1, chol, O.rhf_rdm1(c), 0.01)
    fb_trial = F.force_bias(chol, c, c, c, c)  # <T|L_g|T>/<T|T>, both spins
    return bool(np.allclose(ctx.mf_shifts, 1j * fb_trial, rtol=1e-12, atol=1e-13))


def check_theta_known_answers() -> bool:
    """Three walkers; dt = 0.01, E_shift = 1.5, h0_prop = -2 + 0.7i, e_estimate = -1.

    walker 0: -sqrt(dt) shift_term = 0.5i, ratio = 1.5 e^{0.25i}, fb_term = 0.2 + 0.3i:
      Re exponent = 0.2 + 0.01 (1.5 - 2) = 0.195, theta = 0.5 + 0.25 = 0.75
      (the full exponent's phase would add 0.3 + 0.007 and give cos 1.057 instead of cos 0.75);
    walker 1: -sqrt(dt) shift_term = -0.3 + 2i, ratio = 0.8 e^{i (2 pi - 2.4)}, fb_term = -0.1 - 0.2i:
      Re exponent = -0.3 - 0.1 - 0.005 = -0.405, theta = 2 + 2 pi - 2.4 wraps to -0.4;
    walker 2: as walker 0 with ratio 1.5 e^{2i}: theta = 2.5, cos < 0 -> weight 0 and one node.
    """
    dt = 0.01
    ctx = O.PropCtx(dt=dt, sqrt_dt=0.1, exp_h1_half=None, mf_shifts=None, h0_prop=-2.0 + 0.7j, chol_flat=None,
                    norb=0)
    shift = np.array([-5j, 3.0 - 20j, -5j])
    fbt = np.array([0.2 + 0.3j, -0.1 - 0.2j, 0.2 + 0.3j])
    ratio = np.array([1.5 * np.exp(0.25j), 0.8 * np.exp(1j * (2 * math.pi - 2.4)), 1.5 * np.exp(2.0j)])
    weights = np.array([1.0, 2.0, 1.0])
    w, nodes, pop = O.phaseless_update(weights, np.ones(3, complex), ratio, shift, fbt, 1.5, -1.0, 0, ctx,
                                       O.QmcParams(dt=dt))
    expect = np.array([1.5 * math.exp(0.195) * math.cos(0.75), 2.0 * 0.8 * math.exp(-0.405) * math.cos(0.4), 0.0])
    pop_expect = -1.0 - 0.1 * math.log(expect.sum() / 3.0) / dt
    return bool(np.allclose(w, expect, rtol=1e-12, atol=0) and nodes == 1
                and math.isclose(pop, pop_expect, rel_tol=1e-12))


CHECKS = {
    "importance_sampling_is_unbiased": check_importance_sampling_is_unbiased,
    "weighted_propagator_is_exp_minus_dt_h": check_weighted_propagator_is_exp_minus_dt_h,
    "force_bias_shift_is_optimal": check_force_bias_shift_is_optimal,
    "mean_field_is_trial_force_bias": check_mean_field_is_trial_force_bias,
    "theta_known_answers": check_theta_known_answers,
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_passes(name):
    assert CHECKS[name]()


# ------------------------------------------------------------------------------------- planted mistakes


def _shift_terms_variant(xbar_sign=-1.0, half=0.5, shift_from_x=False):
    def shift_terms(fields, force_bias, ctx):
        fs = O.field_shifts(force_bias, ctx)
        xbar = fields + xbar_sign * fs
        shift_term = np.sum((fields if shift_from_x else xbar) * ctx.mf_shifts, axis=1)
        fb_term = np.sum(fields * fs - half * fs * fs, axis=1)
        return xbar, shift_term, fb_term
    return shift_terms


def _theta_from_full_exponent(weights, overlaps_old, overlaps_new, shift_term, fb_term, pop_shift, e_estimate,
                              node_encounters, ctx, params):
    ratio = overlaps_new / overlaps_old
    exponent = O.importance_exponent(shift_term, fb_term, pop_shift, ctx)
    imp = np.abs(np.exp(exponent) * ratio) * np.cos(np.angle(np.exp(exponent) * ratio))
    imp = np.where(~np.isfinite(imp) | (imp < params.weight_floor), 0.0, imp)
    nodes = int(node_encounters) + int(np.sum(imp <= 0.0))
    imp = np.where(imp > params.weight_cap, 0.0, imp)
    wn = weights * imp
    wn = np.where(wn > params.weight_cap, 0.0, wn)
    return wn, nodes, float(e_estimate - params.pop_control_damping * math.log(max(wn.mean(), 1e-300)) / ctx.dt)


_build = O.build_prop_ctx
_h1_eff = O.h1_eff
_mf_shifts = O.mf_shifts

MUTANTS = {
    "xbar = x + fs": ("shift_terms", _shift_terms_variant(xbar_sign=+1.0)),
    "fb_term without 1/2": ("shift_terms", _shift_terms_variant(half=1.0)),
    "shift_term from x": ("shift_terms", _shift_terms_variant(shift_from_x=True)),
    "fs sign": ("field_shifts", lambda fb, ctx: ctx.sqrt_dt * (1j * fb - ctx.mf_shifts)),
    "fs without mf": ("field_shifts", lambda fb, ctx: -ctx.sqrt_dt * 1j * fb),
    "fs without i": ("field_shifts", lambda fb, ctx: -ctx.sqrt_dt * (fb - ctx.mf_shifts)),
    "exponent +sqrt(dt) shift_term": (
        "importance_exponent", lambda st, ft, e, ctx: ctx.sqrt_dt * st + ft + ctx.dt * (e + ctx.h0_prop)),
    "exponent without h0_prop": ("importance_exponent", lambda st, ft, e, ctx: -ctx.sqrt_dt * st + ft + ctx.dt * e),
    "exponent without fb_term": (
        "importance_exponent", lambda st, ft, e, ctx: -ctx.sqrt_dt * st + ctx.dt * (e + ctx.h0_prop)),
    "theta of the full exponent": ("phaseless_update", _theta_from_full_exponent),
    "theta without the shift term": ("phaseless_theta", lambda st, ratio, ctx: np.angle(ratio)),
    "h0_prop = -h0 + 1/2 sum mf^2": (
        "build_prop_ctx", lambda *a: (lambda c: replace(c, h0_prop=c.h0_prop + np.sum(c.mf_shifts ** 2)))(_build(*a))),
    "h1_eff mean-field sign": (
        "h1_eff", lambda h1, chol, mf: _h1_eff(h1, chol, mf) + 2.0 * np.einsum("g,gik->ik", (1j * mf).real, chol)),
    "h1_eff without 1/2 sum LL": (
        "h1_eff", lambda h1, chol, mf: _h1_eff(h1, chol, mf) + 0.5 * np.einsum("gik,gkj->ij", chol, chol)),
    "mf sign": ("mf_shifts", lambda chol, rdm1: -_mf_shifts(chol, rdm1)),
}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistake_fails_a_check(mutant, monkeypatch):
    attr, fn = MUTANTS[mutant]
    monkeypatch.setattr(O, attr, fn)
    failed = [name for name, chk in CHECKS.items() if not chk()]
    assert failed, f"planted mistake '{mutant}' passes every check"
