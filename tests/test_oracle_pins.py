"""Pins of the CPU oracle (oracle/afqmc_oracle.py) that do not retype its formulas.

* brute force: <T|O|w>/<T|w> from explicit Fock-space matrices (oracle/fock.py)
  for the force bias, local energy and overlap (RHF, RHF-unrestricted, UHF);
* the reference's derivative-of-overlap oracle (trot/meas/auto.py:73-187,
  tolerances of trot tests/test_rhf.py:26-112);
* known answers from trot tests/test_prop_afqmc_step.py:33-82 and
  tests/test_trial_rhf.py:91-112, 115-157;
* closed forms: Taylor polynomial vs matrix powers and expm, QR uniqueness vs
  Gram-Schmidt, comb SR vs a hand-written loop, masking branches, mean-field
  shifts and h1_eff by explicit loops, exp(-dt/2 h1_eff) by eigendecomposition.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

from oracle import afqmc_oracle as O
from oracle import fock as F


def _ham(rng, N, G, h0=0.3):
    a = rng.standard_normal((N, N))
    h1 = 0.5 * (a + a.T)
    b = rng.standard_normal((G, N, N))
    chol = 0.5 * (b + np.swapaxes(b, 1, 2))
    return h0, h1, chol


def _orth(rng, shape, cplx=True):
    x = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    return np.linalg.qr(x)[0]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rhf_against_fock_space(seed):
    rng = np.random.default_rng(seed)
    N, n, G = 5, 2, 4
    h0, h1, chol = _ham(rng, N, G)
    c = _orth(rng, (N, n), cplx=False)
    w = _orth(rng, (3, N, n))
    ops = O.make_rhf_ops(h0, h1, chol, c)
    e, fb, ov = ops.energy(w), ops.force_bias(w), ops.overlap(w)
    for k in range(3):
        assert abs(e[k] - F.local_energy(h0, h1, chol, c, c, w[k], w[k])) < 1e-12 * abs(e[k])
        assert np.allclose(fb[k], F.force_bias(chol, c, c, w[k], w[k]), rtol=1e-12, atol=1e-12)
        assert abs(ov[k] - F.overlap(c, c, w[k], w[k])) < 1e-12 * abs(ov[k])


def test_rhf_unrestricted_walkers_against_fock_space():
    rng = np.random.default_rng(7)
    N, n, G = 5, 2, 3
    h0, h1, chol = _ham(rng, N, G)
    c = _orth(rng, (N, n), cplx=False)
    wu, wd = _orth(rng, (2, N, n)), _orth(rng, (2, N, n))
    ops = O.make_rhf_ops(h0, h1, chol, c, "unrestricted")
    e, fb = ops.energy((wu, wd)), ops.force_bias((wu, wd))
    for k in range(2):
        assert abs(e[k] - F.local_energy(h0, h1, chol, c, c, wu[k], wd[k])) < 1e-12 * abs(e[k])
        assert np.allclose(fb[k], F.force_bias(chol, c, c, wu[k], wd[k]), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("nab", [(2, 1), (3, 2), (1, 1)])
def test_uhf_against_fock_space(nab):
    rng = np.random.default_rng(sum(nab))
    N, G = 5, 4
    na, nb = nab
    h0, h1, chol = _ham(rng, N, G)
    ca, cb = _orth(rng, (N, na), cplx=False), _orth(rng, (N, nb), cplx=False)
    wu, wd = _orth(rng, (2, N, na)), _orth(rng, (2, N, nb))
    ops = O.make_uhf_ops(h0, h1, chol, ca, cb)
    e, fb, ov = ops.energy((wu, wd)), ops.force_bias((wu, wd)), ops.overlap((wu, wd))
    for k in range(2):
        assert abs(e[k] - F.local_energy(h0, h1, chol, ca, cb, wu[k], wd[k])) < 1e-12 * abs(e[k])
        assert np.allclose(fb[k], F.force_bias(chol, ca, cb, wu[k], wd[k]), rtol=1e-12, atol=1e-12)
        assert abs(ov[k] - F.overlap(ca, cb, wu[k], wd[k])) < 1e-12 * abs(ov[k])


def test_against_derivative_of_overlap_oracle():
    """trot/meas/auto.py vs the manual kernels with trot tests/test_rhf.py tolerances."""
    rng = np.random.default_rng(3)
    N, n, G = 5, 2, 7
    h0, h1, chol = _ham(rng, N, G)
    c = _orth(rng, (N, n))
    c = np.real(c) if False else _orth(rng, (N, n), cplx=False)
    w = _orth(rng, (4, N, n))
    ops = O.make_rhf_ops(h0, h1, chol, c)
    assert np.allclose(O.force_bias_auto(w, chol, c), ops.force_bias(w), rtol=1e-7, atol=1e-8)
    assert np.allclose(np.real(O.energy_auto(w, h0, h1, chol, c)), np.real(ops.energy(w)), rtol=5e-3, atol=5e-4)


def test_energy_low_memory_mode_matches_high():
    """trot tests/test_trial_rhf.py:115-157 (there 1e-6 in complex64)."""
    rng = np.random.default_rng(4)
    N, n, G = 6, 3, 5
    h0, h1, chol = _ham(rng, N, G)
    c = _orth(rng, (N, n), cplx=False)
    w = _orth(rng, (3, N, n))
    hi = O.make_rhf_ops(h0, h1, chol, c, memory_mode="high").energy(w)
    lo = O.make_rhf_ops(h0, h1, chol, c, memory_mode="low").energy(w)
    assert np.allclose(lo, hi, rtol=1e-12, atol=1e-12)


def test_energy_is_h0_for_zero_hamiltonian():
    """trot tests/test_trial_rhf.py:91-112."""
    rng = np.random.default_rng(5)
    N, n, G = 5, 2, 4
    c = _orth(rng, (N, n), cplx=False)
    w = c[None] @ (np.eye(n) * (1.1 + 0.2j))[None]
    ops = O.make_rhf_ops(1.234, np.zeros((N, N)), np.zeros((G, N, N)), c)
    assert np.allclose(ops.energy(w), 1.234)
    uops = O.make_rhf_ops(1.234, np.zeros((N, N)), np.zeros((G, N, N)), c, "unrestricted")
    assert np.allclose(uops.energy((w, w)), 1.234)


def test_force_bias_unrestricted_equals_restricted_for_equal_spins():
    """trot tests/test_trial_rhf.py:72-88."""
    rng = np.random.default_rng(6)
    N, n, G = 6, 3, 3
    h0, h1, chol = _ham(rng, N, G)
    c = _orth(rng, (N, n), cplx=False)
    w = c[None] @ (rng.standard_normal((1, n, n)) + 1.5 * np.eye(n))
    r = O.make_rhf_ops(h0, h1, chol, c)
    u = O.make_rhf_ops(h0, h1, chol, c, "unrestricted")
    assert np.allclose(u.force_bias((w, w)), r.force_bias(w), rtol=1e-13)
    assert np.allclose(u.energy((w, w)), r.energy(w), rtol=1e-13)
    assert np.allclose(u.overlap((w, w)), r.overlap(w), rtol=1e-13)


def _dummy_ops(G):
    return O.RhfOps(overlap=lambda w: np.ones(w.shape[0], dtype=complex),
                    force_bias=lambda w: np.zeros((w.shape[0], G), dtype=complex), energy=None)


def test_weight_update_known_answer():
    """trot tests/test_prop_afqmc_step.py:33-82: chol = 0, h0 = 1, dt = 0.2."""
    N, n, W, G = 4, 2, 8, 3
    params = O.QmcParams(dt=0.2, n_exp_terms=4, pop_control_damping=0.1)
    ctx = O.build_prop_ctx(1.0, np.zeros((N, N)), np.zeros((G, N, N)), np.zeros((N, N)), params.dt)
    st = O.PropState(walkers=np.ones((W, N, n), dtype=complex), weights=np.ones(W), overlaps=np.ones(W, complex),
                     pop_control_ene_shift=0.0, e_estimate=0.0)
    out = O.afqmc_step(st, np.random.default_rng(0).standard_normal((W, G)), ctx, _dummy_ops(G), params)
    assert np.allclose(out.weights, np.exp(-0.2))
    assert np.isclose(out.pop_control_ene_shift, 0.1)
    assert out.e_estimate == 0.0


def test_step_equals_manual_trotter():
    """trot tests/test_prop_afqmc_step.py:85-162: with zero force bias and zero rdm1, walkers = trotter(w, x)."""
    rng = np.random.default_rng(42)
    N, n, W, G = 5, 2, 6, 4
    a = rng.standard_normal((N, N))
    h1 = 0.05 * (a + a.T)
    chol = 0.02 * rng.standard_normal((G, N, N))
    params = O.QmcParams(dt=0.1, n_exp_terms=6)
    ctx = O.build_prop_ctx(0.0, h1, chol, np.zeros((N, N)), params.dt)
    w = rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n))
    x = rng.standard_normal((W, G))
    st = O.PropState(walkers=w, weights=np.ones(W), overlaps=np.ones(W, complex), pop_control_ene_shift=0.0,
                     e_estimate=0.0)
    out = O.afqmc_step(st, x, ctx, _dummy_ops(G), params)
    e = scipy.linalg.expm(-0.5 * params.dt * (h1 - 0.5 * np.einsum("gik,gkj->ij", chol, chol)))
    expected = []
    for k in range(W):
        v = np.einsum("g,gij->ij", x[k], chol)
        wk = e @ w[k]
        term, acc = wk, wk
        for j in range(1, 7):
            term = (1j * np.sqrt(params.dt) / j) * (v @ term)
            acc = acc + term
        expected.append(e @ acc)
    assert np.allclose(out.walkers, np.array(expected), rtol=1e-13, atol=1e-14)
    assert np.allclose(out.overlaps, 1.0)


def test_taylor_is_the_truncated_exponential():
    rng = np.random.default_rng(8)
    N, n, K = 7, 3, 6
    v = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    v = 0.1 * (v + v.T)
    w = rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n))
    a = 1j * 0.1
    got = O.taylor_expm_action(a, v, w, K)
    poly = sum(np.linalg.matrix_power(a * v, k) @ w / np.math.factorial(k) for k in range(K + 1)) \
        if hasattr(np, "math") else None
    import math

    poly = sum(np.linalg.matrix_power(a * v, k) @ w / math.factorial(k) for k in range(K + 1))
    assert np.allclose(got, poly, rtol=1e-14, atol=1e-15)
    nrm = np.linalg.norm(a * v, 2)
    bound = nrm ** (K + 1) / math.factorial(K + 1) * math.exp(nrm) * np.linalg.norm(w, 2)
    assert np.linalg.norm(got - scipy.linalg.expm(a * v) @ w, 2) <= bound * 1.0001 + 1e-15


def test_qr_phase_is_gram_schmidt():
    """QR with positive real diag(R) is unique: compare with classical Gram-Schmidt in Python loops."""
    rng = np.random.default_rng(9)
    N, n = 6, 4
    w = rng.standard_normal((2, N, n)) + 1j * rng.standard_normal((2, N, n))
    q, det_r = O.qr_phase(w)
    for b in range(2):
        cols = []
        rdiag = []
        for j in range(n):
            v = w[b][:, j].copy()
            for u in cols:
                v = v - (u.conj() @ w[b][:, j]) * u
            nv = np.sqrt(np.sum(np.abs(v) ** 2))
            cols.append(v / nv)
            rdiag.append(nv)
        assert np.allclose(q[b], np.array(cols).T, atol=1e-12)
        assert np.isclose(det_r[b], np.prod(rdiag))
        assert np.allclose(q[b].conj().T @ q[b], np.eye(n), atol=1e-13)


def test_sr_comb_known_answer_and_loop():
    idx = O.sr_indices(np.array([1.0, 0.0, 0.0, 3.0]), 0.5, 4)
    assert idx.tolist() == [0, 3, 3, 3]
    rng = np.random.default_rng(10)
    wts = np.abs(rng.standard_normal(50)) * (rng.random(50) > 0.2)
    zeta = 0.77
    idx = O.sr_indices(wts, zeta, 50)
    tot = sum(abs(x) for x in wts)
    for i in range(50):
        z = tot * (i + zeta) / 50
        acc, j = 0.0, 0
        while True:
            acc += abs(wts[j])
            if acc >= z or j == 49:
                break
            j += 1
        assert idx[i] == j
    _, wn, _ = O.stochastic_reconfiguration(np.zeros((50, 2, 1)), wts, zeta, "restricted")
    assert np.isclose(wn.sum(), np.abs(wts).sum())


def test_masking_branches_known_answers():
    """Zero shift and bias terms: importance = Re(ratio); floor / node / cap branches by hand."""
    ratio = np.array([2.0, -1.0, 0.5 + 0.5j, 1e-4, 150.0, np.nan, 1j, 60.0])
    weights = np.array([1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 2.0])
    ctx = O.PropCtx(dt=0.01, sqrt_dt=0.1, exp_h1_half=None, mf_shifts=None, h0_prop=0.0, chol_flat=None, norb=0)
    zero = np.zeros(8, dtype=complex)
    w, nodes, pop = O.phaseless_update(weights, np.ones(8, complex), ratio, zero, zero, 0.0, -2.0, 5, ctx,
                                       O.QmcParams(dt=0.01))
    assert np.allclose(w, [2.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0], rtol=1e-15, atol=0)  # 60*2 > cap -> 0
    assert nodes == 5 + 4  # -1, 1e-4 (floored), nan, 1j (cos = 0)
    assert np.isclose(pop, -2.0 - 0.1 * np.log(3.0 / 8) / 0.01)


def test_prop_ctx_closed_forms():
    rng = np.random.default_rng(11)
    N, G, n = 5, 3, 2
    h0, h1, chol = _ham(rng, N, G)
    c = _orth(rng, (N, n), cplx=False)
    dm = O.rhf_rdm1(c)
    ctx = O.build_prop_ctx(h0, h1, chol, dm, 0.05)
    dmt = dm[0] + dm[1]
    for g in range(G):
        s = sum(chol[g, i, j] * dmt[j, i] for i in range(N) for j in range(N))
        assert np.isclose(ctx.mf_shifts[g], 1j * s)
    v0 = np.zeros((N, N))
    for g in range(G):
        for i in range(N):
            for j in range(N):
                v0[i, j] += 0.5 * sum(chol[g, i, k] * chol[g, k, j] for k in range(N))
    he = h1 - v0 - np.einsum("g,gij->ij", (1j * ctx.mf_shifts).real, chol)
    evals, evecs = np.linalg.eigh(he)
    assert np.allclose(ctx.exp_h1_half, evecs @ np.diag(np.exp(-0.025 * evals)) @ evecs.T, atol=1e-13)
    assert np.isclose(ctx.h0_prop, -h0 - 0.5 * np.sum(ctx.mf_shifts ** 2))
    z = O.build_prop_ctx(h0, h1, np.zeros_like(chol), dm, 0.05)
    assert np.allclose(z.mf_shifts, 0) and np.isclose(z.h0_prop, -h0)


def test_block_clamps_and_conserves_weight():
    """blocks.py:236-272: NaN energies -> E_est with weight 0; far energies clamped; SR conserves sum |w|."""
    rng = np.random.default_rng(12)
    N, n, W, G = 4, 2, 6, 3
    h0, h1, chol = _ham(rng, N, G)
    c = _orth(rng, (N, n), cplx=False)
    base = O.make_rhf_ops(h0, h1, chol, c)
    planted = np.array([np.nan, 1e6, 0.0, 0.0, 0.0, 0.0])
    ops = O.RhfOps(overlap=base.overlap, force_bias=base.force_bias,
                   energy=lambda w: base.energy(w) * 0 + planted + np.arange(W) * 0.01)
    params = O.QmcParams(dt=0.01, n_prop_steps=0, shift_ema=0.1)
    ctx = O.build_prop_ctx(h0, h1, chol, O.rhf_rdm1(c), 0.01)
    w0 = c[None] + 0.01 * (rng.standard_normal((W, N, n)) + 0j)
    st = O.PropState(walkers=w0, weights=np.arange(1.0, W + 1), overlaps=base.overlap(w0), pop_control_ene_shift=0.0,
                     e_estimate=0.0)
    st2, obs = O.block(st, [], 0.3, ctx, ops, params)
    e = np.array([0.0, 0.0, 0.02, 0.03, 0.04, 0.05])  # NaN and 1e6 replaced by E_est = 0
    wts = np.array([0.0, 2, 3, 4, 5, 6])
    assert np.isclose(obs["energy"], np.sum(wts * e) / wts.sum())
    assert np.isclose(obs["weight"], wts.sum())
    assert np.isclose(st2.e_estimate, 0.1 * obs["energy"])
    assert np.isclose(st2.weights.sum(), wts.sum())


def test_init_walkers_span_trial():
    rng = np.random.default_rng(13)
    N, n = 6, 3
    c = _orth(rng, (N, n), cplx=False)
    w = O.init_walkers(O.rhf_rdm1(c), (n, n), 4, "restricted")
    assert w.shape == (4, N, n)
    assert np.allclose(np.abs(np.linalg.det(c.T @ w[0])), 1.0)


def test_qr_phase_ill_conditioned():
    """The reference's Householder QR with the phase fix (trot/walkers.py:137-148) stays orthonormal for
    kappa = 1e10 walkers (where Cholesky-QR breaks down): Q^H Q = I, Q R = w, R upper triangular with a real
    positive diagonal whose product is the product of the singular values."""
    rng = np.random.default_rng(45)
    N, n = 30, 6
    u = np.linalg.qr(rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n)))[0]
    v = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    s = np.logspace(0.0, -10.0, n)
    w = ((u * s) @ v.conj().T)[None]
    q, det_r = O.qr_phase(w)
    assert np.max(np.abs(q[0].conj().T @ q[0] - np.eye(n))) < 1e-13
    r = q[0].conj().T @ w[0]
    assert np.max(np.abs(q[0] @ r - w[0])) < 1e-14
    assert np.max(np.abs(np.tril(r, -1))) < 1e-14
    assert np.all(np.diagonal(r).real > 0) and np.max(np.abs(np.diagonal(r).imag)) < 1e-14
    assert abs(det_r[0] - np.prod(s)) < 1e-5 * np.prod(s)
