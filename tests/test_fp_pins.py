"""Pins of the oracle's free-projection functions (oracle/afqmc_oracle.py, free projection section).

* qr_norm against det(w^H w) from numpy's determinant (|det R|^2 = det(w^H w) for w = QR);
* orthogonalize: Q^H Q = I, Q^H w upper triangular with real positive diagonal whose product is
  sqrt(det(w^H w)) per spin;
* multiply_constants: the overlap with any determinant scales by exactly the walker's constant
  (restricted: det(C^T w c^(1/2n))^2 = c det(C^T w)^2; unrestricted: c^(n_a/n) c^(n_b/n) = c);
* the step with zero one- and two-body terms (V = 0, E = I, mf = 0): walkers are scaled by
  exp(dt (ene0 - h0))^(1/2n), the weights after the comb are sum|w D| / W / D[idx] with D from
  numpy's determinant and the comb indices written out by hand;
* the block with zero two-body terms: E_L = h0 for every walker, so the block energy is h0 (or ene0
  when |h0 - ene0| > sqrt(2/dt)) and the block weight is sum w <T|w>.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import afqmc_oracle as O


def _walkers(rng, W, N, n):
    return rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n))


def test_qr_norm_is_gram_determinant():
    rng = np.random.default_rng(0)
    w = _walkers(rng, 5, 9, 4)
    gram_det = np.linalg.det(np.conj(np.swapaxes(w, 1, 2)) @ w).real
    assert np.allclose(O.qr_norm(w, "restricted"), gram_det, rtol=1e-12)
    wd = _walkers(rng, 5, 9, 3)
    gd = np.linalg.det(np.conj(np.swapaxes(wd, 1, 2)) @ wd).real
    assert np.allclose(O.qr_norm((w, wd), "unrestricted"), np.sqrt(gram_det * gd), rtol=1e-12)


def test_orthogonalize_q_and_norm():
    rng = np.random.default_rng(1)
    w = _walkers(rng, 4, 8, 3)
    q, norm = O.orthogonalize(w, "restricted")
    qh = np.conj(np.swapaxes(q, 1, 2))
    assert np.allclose(qh @ q, np.eye(3)[None], atol=1e-12)
    r = qh @ w
    assert np.allclose(np.tril(r, -1), 0.0, atol=1e-12)
    d = np.diagonal(r, axis1=1, axis2=2)
    assert np.allclose(d.imag, 0.0, atol=1e-12) and np.all(d.real > 0)
    assert np.allclose(norm, np.linalg.det(np.conj(np.swapaxes(w, 1, 2)) @ w).real, rtol=1e-12)


def test_multiply_constants_scales_overlap_by_constant():
    rng = np.random.default_rng(2)
    N, n, W = 7, 3, 4
    c_trial = np.linalg.qr(rng.standard_normal((N, n)))[0]
    w = _walkers(rng, W, N, n)
    const = np.exp(rng.standard_normal(W) + 1j * rng.uniform(-3, 3, W))
    ov = O.overlap_r(w, c_trial)
    assert np.allclose(O.overlap_r(O.multiply_constants(w, const, "restricted"), c_trial), const * ov, rtol=1e-12)
    cb = np.linalg.qr(rng.standard_normal((N, n - 1)))[0]
    wd = _walkers(rng, W, N, n - 1)
    ovu = O.overlap_u((w, wd), c_trial, cb)
    scaled = O.multiply_constants((w, wd), const, "unrestricted")
    assert np.allclose(O.overlap_u(scaled, c_trial, cb), const * ovu, rtol=1e-12)


def test_fp_step_without_interactions():
    rng = np.random.default_rng(3)
    N, n, G, W = 6, 2, 3, 5
    h0, ene0, dt = 0.7, -0.2, 0.05
    chol = np.zeros((G, N, N))
    c_trial = np.linalg.qr(rng.standard_normal((N, n)))[0]
    ctx = O.build_prop_ctx_fp(h0, np.zeros((N, N)), chol, O.rhf_rdm1(c_trial), dt, ene0)
    w = _walkers(rng, W, N, n)
    weights = rng.uniform(0.5, 1.5, W)
    st = O.PropState(walkers=w, weights=weights, overlaps=np.ones(W, complex), pop_control_ene_shift=0.0,
                     e_estimate=0.0)
    zeta = 0.37
    out = O.afqmc_step_fp(st, rng.standard_normal((W, G)), zeta, ctx, "restricted")
    c = np.exp(dt * (ene0 - h0))
    scaled = w * c ** (1.0 / (2 * n))
    D = np.linalg.det(np.conj(np.swapaxes(scaled, 1, 2)) @ scaled).real
    wn = weights * D
    cw = np.cumsum(np.abs(wn))
    idx = [min(int(np.sum(cw < cw[-1] * (i + zeta) / W)), W - 1) for i in range(W)]
    assert np.allclose(out.walkers, scaled[idx], rtol=1e-12, atol=1e-14)
    assert np.allclose(out.weights, (cw[-1] / W) / D[idx], rtol=1e-12)
    assert np.allclose(D, c * np.linalg.det(np.conj(np.swapaxes(w, 1, 2)) @ w).real, rtol=1e-12)


@pytest.mark.parametrize("ene0", [0.1, 40.0])
def test_fp_block_without_two_body(ene0):
    rng = np.random.default_rng(4)
    N, n, G, W = 6, 2, 2, 4
    h0, dt = 0.3, 0.05
    a = rng.standard_normal((N, N))
    h1 = 0.1 * (a + a.T)
    chol = np.zeros((G, N, N))
    c_trial = np.linalg.eigh(h1)[1][:, :n]
    ops = O.make_rhf_ops(h0, np.zeros((N, N)), chol, c_trial)  # E_L = h0 exactly
    ctx = O.build_prop_ctx_fp(h0, h1, chol, O.rhf_rdm1(c_trial), dt, ene0)
    params = O.QmcParamsFp(dt=dt, n_prop_steps=4, n_qr_blocks=2, ene0=ene0, n_walkers=W)
    w = c_trial[None] + 0.1 * _walkers(rng, W, N, n)
    st = O.PropState(walkers=w, weights=np.ones(W), overlaps=ops.overlap(w), pop_control_ene_shift=0.0,
                     e_estimate=0.0)
    fields = [rng.standard_normal((W, G)) for _ in range(4)]
    st, obs = O.block_fp(st, fields, [0.2, 0.5, 0.8, 0.1], ctx, ops, params)
    expect = h0 if abs(h0 - ene0) <= np.sqrt(2.0 / dt) else ene0
    assert np.isclose(obs["energy"], expect, rtol=1e-12)
    assert np.isclose(obs["weight"], np.sum(st.weights * ops.overlap(st.walkers)), rtol=1e-12)
    assert np.isclose(obs["abs_overlap"], np.sum(np.abs(st.overlaps)), rtol=1e-12)
