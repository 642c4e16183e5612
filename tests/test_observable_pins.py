"""Pins of the block observables and the block statistics (CPU).

Observables (trot/meas/rhf.py:191-261, trot/meas/uhf.py:97-160): the oracle's Green's-function formulas
for the spin-resolved one-body density matrix and the density correlations are checked against explicit
second-quantised operators on the Fock space of tiny systems (``oracle/fock.py`` builds the strings and
the a+_p a_q matrices from minors, without Green's functions):

    rdm1[s][a, b] = <T| a+_{b s} a_{a s} |w> / <T|w>          (the reference's index order)
    uu[i, j] = <n_{i up} n_{j up}>, ud[i, j] = <n_{i up} n_{j dn}>, dd[i, j] = <n_{i dn} n_{j dn}>

Statistics (trot/stat_utils.py:12-313): closed forms that a dropped factor or a shifted index breaks:
with equal weights the leave-one-block-out SE at block size B is the textbook standard error of the
block means; the jackknife of a ratio with unit denominators is the textbook SE of the mean; the
rescaled LOO samples have mean mu and spread SE* sqrt((G - 1) / G); and hand-worked plateau cases.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import afqmc_oracle as orc
from oracle import fock
from trot_amd import stats


def _cplx(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _spin_ops(norb, nu, nd):
    eu, ed = fock._one_spin_epq(norb, nu), fock._one_spin_epq(norb, nd)
    iu, idn = np.eye(eu.shape[-1]), np.eye(ed.shape[-1])
    dim = eu.shape[-1] * ed.shape[-1]
    up = np.einsum("pqab,cd->pqacbd", eu, idn).reshape(norb, norb, dim, dim)
    dn = np.einsum("ab,pqcd->pqacbd", iu, ed).reshape(norb, norb, dim, dim)
    return up, dn


def _mixed(op, bra, ket):
    return complex(bra.conj() @ op @ ket / (bra.conj() @ ket))


def _brute(ca, cb, wu, wd):
    norb = ca.shape[0]
    fs = fock.FockSpace(norb, (ca.shape[1], cb.shape[1]))
    bra, ket = fs.state(ca, cb), fs.state(wu, wd)
    up, dn = _spin_ops(norb, ca.shape[1], cb.shape[1])
    dm = np.array([[[_mixed(e[b, a], bra, ket) for b in range(norb)] for a in range(norb)] for e in (up, dn)])
    corr = np.zeros((3, norb, norb), complex)
    for i in range(norb):
        for j in range(norb):
            corr[0, i, j] = _mixed(up[i, i] @ up[j, j], bra, ket)
            corr[1, i, j] = _mixed(up[i, i] @ dn[j, j], bra, ket)
            corr[2, i, j] = _mixed(dn[i, i] @ dn[j, j], bra, ket)
    return dm, corr


def test_rdm1_and_density_corr_restricted_match_fock_space():
    rng = np.random.default_rng(7)
    N, n = 4, 2
    c = np.linalg.qr(rng.standard_normal((N, N)))[0][:, :n]
    w = _cplx(rng, 2, N, n)
    dm = orc.rdm1_rw(w, c)
    corr = orc.density_corr_from_rdm1(dm)
    for k in range(2):
        b_dm, b_corr = _brute(c, c, w[k], w[k])
        assert np.abs(dm[k] - b_dm).max() < 1e-11
        assert np.abs(corr[k] - b_corr).max() < 1e-11


def test_rdm1_and_density_corr_unrestricted_match_fock_space():
    rng = np.random.default_rng(8)
    N, na, nb = 4, 2, 1
    ca = np.linalg.qr(rng.standard_normal((N, N)))[0][:, :na]
    cb = np.linalg.qr(rng.standard_normal((N, N)))[0][:, :nb]
    wu, wd = _cplx(rng, 2, N, na), _cplx(rng, 2, N, nb)
    dm = orc.rdm1_uw((wu, wd), ca, cb)
    corr = orc.density_corr_from_rdm1(dm)
    for k in range(2):
        b_dm, b_corr = _brute(ca, cb, wu[k], wd[k])
        assert np.abs(dm[k] - b_dm).max() < 1e-11
        assert np.abs(corr[k] - b_corr).max() < 1e-11
    # trace of each spin's matrix = that spin's electron count
    tr = np.trace(dm, axis1=2, axis2=3)
    assert np.allclose(tr, [[na, nb]] * 2, atol=1e-12)


def test_weighted_block_mean_closed_forms():
    x = np.arange(12.0).reshape(3, 2, 2) + 1j
    w = np.array([1.0, 0.0, 3.0])
    assert np.allclose(orc.weighted_block_mean(x, w), (x[0] + 3 * x[2]) / 4)
    assert np.array_equal(orc.weighted_block_mean(x, np.zeros(3)), np.zeros((2, 2), complex))


def test_block_grid_adds_the_largest_size_that_leaves_five_blocks():
    assert stats.default_block_grid(100) == [1, 2, 3, 4, 5]
    assert stats.default_block_grid(50) == [1, 2]
    # 30 samples: 2 leaves 15 < 20 blocks, but >= 5, so it is appended
    assert stats.default_block_grid(30) == [1, 2]
    assert stats.default_block_grid(39) == [1, 2]
    assert stats.default_block_grid(9) == []  # 9 // 1 < 20 blocks and 9 // 2 = 4 < 5 blocks
    g = stats.default_block_grid(5000)
    assert g[0] == 1 and g[-1] == 250 and all(5000 // b >= 20 for b in g) and g == sorted(set(g))


@pytest.mark.parametrize("b", [1, 2, 5, 10])
def test_loo_se_with_equal_weights_is_the_textbook_standard_error(b):
    rng = np.random.default_rng(b)
    e = rng.standard_normal(400)
    out = stats.blocking_analysis_ratio(e, np.ones_like(e), block_grid=[b], min_blocks=1)
    means = e.reshape(-1, b).mean(axis=1)
    assert out["B_star"] == b
    assert abs(out["se_star"] - means.std(ddof=1) / np.sqrt(means.size)) < 1e-13
    assert abs(out["mu"] - e.mean()) < 1e-14
    g = means.size
    est = out["estimator_scale_samples"]
    assert abs(est.mean() - out["mu"]) < 1e-13
    assert abs(est.std() - out["se_star"] * np.sqrt((g - 1) / g)) < 1e-13
    lo, hi = out["ci95_star"]
    assert abs(hi - lo - 2 * 1.96 * out["se_star"]) < 1e-13


def test_ratio_estimator_weights_and_exact_bias():
    rng = np.random.default_rng(3)
    e = rng.standard_normal(300) + 2.0
    w = rng.uniform(0.5, 1.5, 300)
    out = stats.blocking_analysis_ratio(e, w, exact=2.0)
    assert abs(out["mu"] - np.sum(w * e) / np.sum(w)) < 1e-13
    assert abs(out["bias"] - (out["mu"] - 2.0)) < 1e-15
    assert abs(out["z_score"] - out["bias"] / out["se_star"]) < 1e-12
    assert list(out["block_sizes"]) == stats.default_block_grid(300)
    assert list(out["n_blocks"]) == [300 // b for b in out["block_sizes"]]
    # fewer than 5 blocks at every size: no curve
    none = stats.blocking_analysis_ratio(e[:4], w[:4])
    assert none["B_star"] is None and none["se_star"] is None and none["ci95_star"] == (None, None)


def test_pick_plateau_hand_worked_cases():
    bs = np.array([1, 2, 4, 8, 16, 32, 64])
    gs = np.array([640, 320, 160, 80, 40, 20, 10])
    # rises past 1.2 SE(1) at B=4 and stays within 3% for the next 3 sizes
    se = np.array([1.0, 1.1, 1.3, 1.31, 1.32, 1.33, 1.5])
    assert stats._pick_plateau(bs, se, gs) == (4, 1.3, 160)
    # B=64 is dropped (G < 20), so windows of k + 1 = 4 sizes start at i = 0, 1, 2 only; the one at B=4
    # has risen but steps 10% to B=8: no plateau, and the 95% rule (of 1.45) picks B=8
    se = np.array([1.0, 1.1, 1.3, 1.43, 1.44, 1.45, 9.0])
    bs2 = bs[:6]
    assert stats._pick_plateau(bs, se, gs) == (8, 1.43, 80)
    # no plateau, steadily rising: first size within 95% of the largest SE
    se = np.array([1.0, 1.5, 2.0, 3.0, 4.0, 4.1, 9.0])
    assert stats._pick_plateau(bs, se, gs) == (16, 4.0, 40)
    # no size keeps min_blocks blocks: the first size of the curve
    assert stats._pick_plateau(bs2, np.ones(6), np.full(6, 10)) == (1, 1.0, 10)
    # flat but never risen by 20%: the 95% rule picks the first size
    assert stats._pick_plateau(bs, np.ones(7), gs) == (1, 1.0, 640)


def test_jackknife_ratio_with_unit_denominators_is_the_textbook_error():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((50, 3, 2)) + 1j * rng.standard_normal((50, 3, 2))
    mean, sigma = stats.jackknife_ratios(x, np.ones(50))
    assert np.allclose(mean, x.real.mean(axis=0), atol=1e-14)
    assert np.allclose(sigma, x.real.std(axis=0, ddof=1) / np.sqrt(50), atol=1e-14)


def test_rebin_observable_sums_weighted_super_blocks():
    obs = np.arange(14.0).reshape(7, 2)
    w = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0])
    num, den = stats.rebin_observable(obs, w, 3)
    assert np.array_equal(den, [6.0, 15.0])  # the 7th block is dropped
    assert np.array_equal(num[0], 1 * obs[0] + 2 * obs[1] + 3 * obs[2])
    assert np.array_equal(num[1], 4 * obs[3] + 5 * obs[4] + 6 * obs[5])


def test_reject_outliers_drops_the_far_block_only():
    e = np.array([-1.0, -1.01, -0.99, -1.02, -0.98, 50.0])
    data = np.column_stack([e, np.ones(6)])
    kept, mask = stats.reject_outliers(data, obs=0)
    assert mask.tolist() == [True] * 5 + [False]
    assert kept.shape == (5, 2)
