"""Pins of the hydrogen-chain staging (oracle/h_chain.py) against independent routes (CPU).

* one-electron integrals of the contracted STO-6G functions against 2-D quadrature in cylindrical
  coordinates (overlap, kinetic energy as 1/2 int grad.grad, nuclear attraction with the 1/r factor
  integrated directly), and the normalization against a radial quadrature;
* electron repulsion against the Gaussian transform 1/r = 2/sqrt(pi) int_0^inf exp(-u^2 r^2) du,
  integrated numerically over u (no Boys function);
* the staged H12 system: converged SCF (FDS = SDF), Cholesky residual within the cut, the identity
  trial of the MO basis, and the RHF energy recovered from h1 and the MO Cholesky vectors.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy import integrate

from oracle import h_chain as H


def _phi(a, c, center, x, y, z):
    r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    return np.sum(c * np.exp(-a * r2))


def _grad_phi(a, c, center, x, y, z):
    d = np.array([x - center[0], y - center[1], z - center[2]])
    r2 = np.sum(d ** 2)
    g = np.sum(c * (-2.0 * a) * np.exp(-a * r2))
    return g * d


def test_contraction_normalized_radially():
    a, c = H._contraction()
    val, _ = integrate.quad(lambda r: 4 * np.pi * r * r * np.sum(c * np.exp(-a * r * r)) ** 2, 0, 40, limit=200)
    assert abs(val - 1.0) < 1e-9


def test_hydrogen_atom_radial():
    """T and V of one H atom against radial quadrature: T = 1/2 int |phi'|^2, V = -int phi^2 / r."""
    ints = H.integrals(np.zeros((1, 3)))
    a, c = H._contraction()
    dphi = lambda r: np.sum(c * (-2 * a * r) * np.exp(-a * r * r))  # noqa: E731
    phi = lambda r: np.sum(c * np.exp(-a * r * r))  # noqa: E731
    T, _ = integrate.quad(lambda r: 0.5 * 4 * np.pi * r * r * dphi(r) ** 2, 0, 40, limit=200, epsabs=1e-13)
    V, _ = integrate.quad(lambda r: -4 * np.pi * r * phi(r) ** 2, 0, 40, limit=200, epsabs=1e-13)
    assert abs(ints.T[0, 0] - T) < 1e-9 and abs(ints.V[0, 0] - V) < 1e-9
    assert ints.enuc == 0.0


@pytest.mark.parametrize("dist", [1.8])
def test_one_electron_integrals_quadrature(dist):
    coords = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, dist]])
    ints = H.integrals(coords)
    a, c = H._contraction()
    A, B = coords

    def cyl(f):  # int f dV over the axis-symmetric integrand, rho in [0, 12], z in [-12, dist + 12]
        v, _ = integrate.dblquad(lambda rho, z: 2 * np.pi * rho * f(rho, 0.0, z), -12, dist + 12, 0, 12,
                                 epsabs=1e-10, epsrel=1e-9)
        return v

    S = cyl(lambda x, y, z: _phi(a, c, A, x, y, z) * _phi(a, c, B, x, y, z))
    T = cyl(lambda x, y, z: 0.5 * float(_grad_phi(a, c, A, x, y, z) @ _grad_phi(a, c, B, x, y, z)))
    V = -cyl(lambda x, y, z: _phi(a, c, A, x, y, z) * _phi(a, c, B, x, y, z) * (
        1.0 / max(np.sqrt(x * x + y * y + z * z), 1e-300) + 1.0 / max(np.sqrt(x * x + y * y + (z - dist) ** 2), 1e-300)))
    assert abs(ints.S[0, 1] - S) < 1e-8
    assert abs(ints.T[0, 1] - T) < 1e-7
    assert abs(ints.V[0, 1] - V) < 1e-6


def _eri_transform(a, c, A, B, Cc, D):
    """(AB|CD) over contracted s functions via the Gaussian transform of 1/r12, summed over primitives."""
    tot = 0.0
    for ai, ci in zip(a, c):
        for bj, cj in zip(a, c):
            p = ai + bj
            P = (ai * A + bj * B) / p
            Kab = np.exp(-ai * bj / p * np.sum((A - B) ** 2))
            for ak, ck in zip(a, c):
                for bl, cl in zip(a, c):
                    q = ak + bl
                    Q = (ak * Cc + bl * D) / q
                    Kcd = np.exp(-ak * bl / q * np.sum((Cc - D) ** 2))
                    R2 = np.sum((P - Q) ** 2)

                    def f(u):
                        den = p * q + u * u * (p + q)
                        return np.pi ** 3 * den ** -1.5 * np.exp(-p * q * u * u * R2 / den)

                    val, _ = integrate.quad(f, 0, np.inf, epsabs=1e-13, epsrel=1e-12, limit=200)
                    tot += ci * cj * ck * cl * Kab * Kcd * 2.0 / np.sqrt(np.pi) * val
    return tot


@pytest.mark.parametrize("idx", [(0, 0, 0, 0), (0, 1, 0, 1), (0, 1, 1, 2)])
def test_electron_repulsion_transform(idx):
    coords = H.chain(3, 1.8)
    ints = H.integrals(coords)
    a, c = H._contraction()
    i, j, k, l = idx
    ref = _eri_transform(a, c, coords[i], coords[j], coords[k], coords[l])
    assert abs(ints.eri[i, j, k, l] - ref) < 1e-9 * max(1.0, abs(ref))


def test_h12_staging_invariants():
    st = H.stage_rhf(H.chain(12, 1.8), 12)
    ints = st.ao
    # nuclear repulsion of a uniform chain
    assert abs(st.h0 - sum((12 - d) / (1.8 * d) for d in range(1, 12))) < 1e-12
    # converged SCF: the Fock matrix commutes with the density in the S metric
    Cocc = st.mo_coeff[:, :6]
    D = Cocc @ Cocc.T
    F = ints.T + ints.V + 2 * np.einsum("pqrs,rs->pq", ints.eri, D) - np.einsum("prqs,rs->pq", ints.eri, D)
    assert np.abs(F @ D @ ints.S - ints.S @ D @ F).max() < 1e-7
    # MO orthonormality and the trial = identity columns (trot/staging.py:1153-1163)
    assert np.abs(st.mo_coeff.T @ ints.S @ st.mo_coeff - np.eye(12)).max() < 1e-12
    assert np.abs(st.trial - np.eye(12)[:, :6]).max() < 1e-12
    # pivoted Cholesky: every residual diagonal within the cut, hence every entry (Cauchy-Schwarz)
    L = H.ao_cholesky(ints.eri, 1e-5)
    res = ints.eri.reshape(144, 144) - np.einsum("gi,gj->ij", L.reshape(-1, 144), L.reshape(-1, 144))
    assert np.diag(res).max() <= 1e-5 and np.abs(res).max() <= 1e-5
    assert st.chol.shape == (L.shape[0], 12, 12)
    # the RHF energy from h1 and the MO Cholesky vectors, to within the decomposition error
    G = st.trial @ st.trial.T
    J = np.einsum("gpq,pq->g", st.chol, G)
    e = st.h0 + 2 * np.sum(G * st.h1) + 2 * np.sum(J ** 2) - np.einsum("gpq,qr,grs,sp->", st.chol, G, st.chol, G)
    assert abs(e - st.e_rhf) < 1e-5
    # symmetric MO Cholesky vectors (the sym-block VHS path applies)
    assert np.abs(st.chol - np.swapaxes(st.chol, 1, 2)).max() < 1e-13
