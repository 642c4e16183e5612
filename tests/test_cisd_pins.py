"""Pins of the CISD oracle (oracle/cisd_oracle.py) that do not restate its formulas.

* brute force: |T> = (1 + sum_ia c_ia E_ai + 1/2 sum_iajb c_iajb E_ai E_bj)|HF> built from explicit
  Fock-space matrices (oracle/fock.py), then <T|w>, <T|L_g|w>/<T|w>, <T|H|w>/<T|w> and
  <T|a+_p a_q|w>/<T|w> (one spin) against the oracle's overlap, force bias, energy and 1-RDM,
  with and without frozen core / outer orbitals;
* the force bias as the derivative of ln <T|(1 + x L_g)|w> at x = 0 (trot/meas/auto.py:73-96);
* ci1 = ci2 = 0 reduces every function to the RHF oracle with the trial C = I[:, :nocc];
* the reference's invariances (tests/test_cisd.py:106-157): an active-space trial equals its
  zero-padded full-space twin; slice_trial_level equals zeroing the dropped amplitudes.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import afqmc_oracle as O
from oracle import cisd_oracle as C
from oracle import fock as F


def _ham(rng, N, G):
    a = rng.standard_normal((N, N))
    b = rng.standard_normal((G, N, N))
    return float(rng.standard_normal()), 0.5 * (a + a.T), 0.5 * (b + np.swapaxes(b, 1, 2))


def _trial(rng, nocc, nvir, core=0, outer=0, s1=0.05, s2=0.02):
    ci1 = s1 * rng.standard_normal((nocc, nvir))
    ci2 = s2 * rng.standard_normal((nocc, nvir, nocc, nvir))
    ci2 = 0.5 * (ci2 + ci2.transpose(2, 3, 0, 1))
    return C.CisdTrial(ci1, ci2, core, outer)


def _walkers(rng, W, N, n, mix=0.25):
    """Near the reference determinant [I; 0], then QR (trot/testing.py:85-102)."""
    out = []
    for _ in range(W):
        w = np.zeros((N, n), complex)
        w[:n] = np.eye(n)
        w = w + mix * (rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n)))
        out.append(np.linalg.qr(w)[0])
    return np.array(out)


def _bra(fs, t: C.CisdTrial):
    N, nf, core = t.norb, t.nocc_full, t.nocc_t_core
    eye = np.eye(N)
    hf = fs.state(eye[:, :nf], eye[:, :nf])
    v = hf.copy()
    single = {}
    for i in range(t.nocc):
        for a in range(t.nvir):
            single[i, a] = fs.epq[nf + a, core + i] @ hf
            v = v + t.ci1[i, a] * single[i, a]
    for (i, a), s_ia in single.items():
        for j in range(t.nocc):
            for b in range(t.nvir):
                v = v + 0.5 * t.ci2[i, a, j, b] * (fs.epq[nf + a, core + i] @ single[j, b])
    return v


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("N,nocc,core,outer", [(5, 2, 0, 0), (6, 3, 0, 0), (6, 3, 1, 1), (6, 2, 1, 2)])
def test_against_fock_space(N, nocc, core, outer):
    rng = np.random.default_rng(11 + N + 3 * core + outer)
    G = 4
    h0, h1, chol = _ham(rng, N, G)
    t = _trial(rng, nocc - core, N - nocc - outer, core, outer, s1=0.2, s2=0.1)
    w = _walkers(rng, 3, N, nocc)
    fs = F.FockSpace(N, (nocc, nocc))
    bra = _bra(fs, t)
    ctx = C.build_meas_ctx(chol, t)
    ov = C.overlap(w, t)
    fb = C.force_bias(w, chol, ctx, t)
    en = C.energy(w, h0, h1, chol, ctx, t)
    dm = C.rdm1(w, t)
    ham = fs.hamiltonian(h0, h1, chol)
    for k in range(w.shape[0]):
        ket = fs.state(w[k], w[k])
        assert abs(ov[k] - bra.conj() @ ket) <= 1e-12 * abs(ov[k])
        fb_bf = np.array([fs.mixed(fs.one_body(lg), bra, ket) for lg in chol])
        assert _rel(fb[k], fb_bf) < 1e-10
        assert abs(en[k] - fs.mixed(ham, bra, ket)) < 1e-10 * abs(en[k])
        gam = np.array([[0.5 * fs.mixed(fs.epq[p, q], bra, ket) for p in range(N)] for q in range(N)])
        # the reference returns G0 - M / overlap (trot/meas/cisd.py:790-818) where the exact mixed
        # estimator is G0 - M^T / overlap: its CI correction is transposed against its HF part
        g0 = np.zeros((N, N), complex)
        g0[:, :nocc] = C.greens(w[k:k + 1], nocc)[0].T
        assert _rel(g0 - (g0 - dm[k, 0]).T, gam) < 1e-10 and np.array_equal(dm[k, 0], dm[k, 1])
        h = h1 + h1.T  # a symmetric one-body operator sees no difference
        assert abs(np.sum(h * dm[k, 0]) - np.sum(h * gam)) < 1e-10 * np.abs(np.sum(h * gam))


def test_force_bias_is_the_overlap_derivative():
    rng = np.random.default_rng(5)
    N, nocc, G = 10, 4, 6
    _, _, chol = _ham(rng, N, G)
    t = _trial(rng, nocc, N - nocc)
    w = _walkers(rng, 2, N, nocc)
    fb = C.force_bias(w, chol, C.build_meas_ctx(chol, t), t)
    auto = C.force_bias_auto(w, chol, lambda x: C.overlap(x, t), eps=1e-4)
    assert _rel(fb, auto) < 1e-8


def test_zero_amplitudes_reduce_to_rhf():
    rng = np.random.default_rng(8)
    N, nocc, G = 9, 3, 5
    h0, h1, chol = _ham(rng, N, G)
    t = C.CisdTrial(np.zeros((nocc, N - nocc)), np.zeros((nocc, N - nocc, nocc, N - nocc)))
    c = np.eye(N)[:, :nocc]
    w = _walkers(rng, 4, N, nocc)
    ctx = C.build_meas_ctx(chol, t)
    rctx = O.build_meas_ctx_rhf(h1, chol, c)
    assert _rel(C.overlap(w, t), O.overlap_r(w, c)) < 1e-13
    assert _rel(C.force_bias(w, chol, ctx, t), O.force_bias_rw(w, rctx, c)) < 1e-13
    assert _rel(C.energy(w, h0, h1, chol, ctx, t), O.energy_rw(w, h0, rctx, c)) < 1e-13
    g = C.greens(w, nocc)
    dm = C.rdm1(w, t)
    assert np.allclose(dm[:, 0, :, :nocc], np.swapaxes(g, 1, 2), rtol=1e-14, atol=1e-14)
    assert not np.any(dm[:, 0, :, nocc:])


@pytest.mark.parametrize("core,outer", [(0, 0), (1, 2)])
def test_active_space_equals_zero_padded_full_space(core, outer):
    """tests/test_cisd.py:106-157 of the reference."""
    rng = np.random.default_rng(321 + core)
    nocc_full, nvir_full, G = 4, 6, 8
    na, va = nocc_full - core, nvir_full - outer
    t = _trial(rng, na, va, core, outer)
    c1 = np.zeros((nocc_full, nvir_full))
    c1[core:, :va] = t.ci1
    c2 = np.zeros((nocc_full, nvir_full, nocc_full, nvir_full))
    c2[core:, :va, core:, :va] = t.ci2
    tf = C.CisdTrial(c1, c2)
    N = nocc_full + nvir_full
    h0, h1, chol = _ham(rng, N, G)
    w = _walkers(rng, 2, N, nocc_full)
    ca, cf = C.build_meas_ctx(chol, t), C.build_meas_ctx(chol, tf)
    assert _rel(C.overlap(w, t), C.overlap(w, tf)) < 1e-12
    assert _rel(C.force_bias(w, chol, ca, t), C.force_bias(w, chol, cf, tf)) < 1e-12
    assert _rel(C.energy(w, h0, h1, chol, ca, t), C.energy(w, h0, h1, chol, cf, tf)) < 1e-12
    assert _rel(C.rdm1(w, t), C.rdm1(w, tf)) < 1e-12


def test_slice_trial_level_drops_amplitudes():
    rng = np.random.default_rng(4)
    t = _trial(rng, 3, 5)
    s = C.slice_trial_level(t, 2)
    assert s.nvir == 2 and s.nvir_t_outer == 3 and s.norb == t.norb
    z1 = t.ci1.copy()
    z1[:, 2:] = 0
    z2 = t.ci2.copy()
    z2[:, 2:] = 0
    z2[:, :, :, 2:] = 0
    w = _walkers(rng, 2, t.norb, 3)
    assert _rel(C.overlap(w, s), C.overlap(w, C.CisdTrial(z1, z2))) < 1e-13
    assert C.slice_trial_level(t, None) is t
