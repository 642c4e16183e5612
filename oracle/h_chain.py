"""Hydrogen-chain Hamiltonians in STO-6G: the molecular parity case (test infrastructure).

Restates the reference's staging of an RHF molecule without PySCF, which the reference calls for
the integrals and the SCF and which is not installed:

* ``stage_rhf`` follows ``_stage_ham_input`` (trot/staging.py:882-975): h0 = nuclear repulsion,
  h1 = C^T hcore C, AO Cholesky vectors rotated to the MO basis (``_rotate_chol_to_mo``,
  staging.py:349-367), no frozen core;
* ``ao_cholesky`` follows the pivoted ("modified") Cholesky of ``chunked_cholesky``
  (staging.py:234-346): pivot on the largest residual diagonal of the full nao^2 ERI matrix, stop
  when it falls to ``max_error``, the last vector computed is dropped;
* the trial is ``_mf_coeff_helper`` (staging.py:1153-1163) of C^T S C, i.e. the identity's first
  nocc columns up to rounding.

Integrals over contracted s Gaussians are the textbook closed forms (overlap, kinetic, nuclear
attraction and repulsion through the Boys function F0); ``tests/test_oracle_pins.py`` pins each of
them against an independent quadrature.  Only tests import this module.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.special import erf

# STO-6G hydrogen 1s (zeta = 1.24): (exponent, contraction coefficient of the normalized primitive).
# Source: W. J. Hehre, R. F. Stewart and J. A. Pople, J. Chem. Phys. 51, 2657 (1969), the least-squares
# six-Gaussian fit of a 1s Slater function, exponents scaled by zeta^2; the values as distributed by the
# Basis Set Exchange (basis "STO-6G", element H).
STO6G_H = (
    (35.52322122, 0.00916359628),
    (6.513143725, 0.04936149294),
    (1.822142904, 0.16853830490),
    (0.625955266, 0.37056279970),
    (0.243076747, 0.41649152980),
    (0.100112428, 0.13033408410),
)


def boys0(t):
    """F0(t) = int_0^1 exp(-t s^2) ds, with the series near t = 0."""
    t = np.asarray(t, dtype=np.float64)
    out = np.empty_like(t)
    small = t < 1e-8
    ts = np.where(small, 1.0, t)
    out[...] = 0.5 * np.sqrt(np.pi / ts) * erf(np.sqrt(ts))
    out[small] = 1.0 - t[small] / 3.0
    return out


def chain(n_atoms: int = 12, bond: float = 1.8) -> np.ndarray:
    """Atom positions (bohr) of a linear chain along z."""
    return np.stack([np.zeros(n_atoms), np.zeros(n_atoms), bond * np.arange(n_atoms)], axis=1)


def _contraction(basis=STO6G_H):
    """Exponents and coefficients of unnormalized primitives, the contraction normalized to 1."""
    a = np.array([e for e, _ in basis])
    c = np.array([c for _, c in basis]) * (2.0 * a / np.pi) ** 0.75
    s = np.sum(c[:, None] * c[None, :] * (np.pi / (a[:, None] + a[None, :])) ** 1.5)
    return a, c / np.sqrt(s)


@dataclass
class Integrals:
    S: np.ndarray  # (nao, nao)
    T: np.ndarray
    V: np.ndarray
    eri: np.ndarray  # (nao, nao, nao, nao), chemists' order (ij|kl)
    enuc: float


def integrals(coords: np.ndarray, charges=None, basis=STO6G_H) -> Integrals:
    """One- and two-electron integrals of one contracted s function per atom."""
    nat = coords.shape[0]
    Z = np.ones(nat) if charges is None else np.asarray(charges, dtype=np.float64)
    a, c = _contraction(basis)
    npr = a.size
    # primitive pairs over all atom pairs: (i, j, pa, pb)
    A = coords[:, None, None, None, :]
    B = coords[None, :, None, None, :]
    ea = a[None, None, :, None]
    eb = a[None, None, None, :]
    p = ea + eb
    mu = ea * eb / p
    AB2 = np.sum((A - B) ** 2, axis=-1)
    K = np.exp(-mu * AB2)
    P = (ea[..., None] * A + eb[..., None] * B) / p[..., None]
    cc = c[None, None, :, None] * c[None, None, None, :]
    s_prim = (np.pi / p) ** 1.5 * K
    S = np.sum(cc * s_prim, axis=(2, 3))
    T = np.sum(cc * mu * (3.0 - 2.0 * mu * AB2) * s_prim, axis=(2, 3))
    V = np.zeros((nat, nat))
    for C, z in zip(coords, Z):
        PC2 = np.sum((P - C) ** 2, axis=-1)
        V -= z * np.sum(cc * (2.0 * np.pi / p) * K * boys0(p * PC2), axis=(2, 3))
    # two-electron: pair index ij over all (i, j), primitive pair index ab
    pp = np.broadcast_to(p, (nat, nat, npr, npr)).reshape(nat * nat, npr * npr)
    KK = (cc * K).reshape(nat * nat, npr * npr)
    PP = P.reshape(nat * nat, npr * npr, 3)
    eri = np.empty((nat * nat, nat * nat))
    for ij in range(nat * nat):
        q = pp[ij][None, :, None]  # (1, ab, 1)
        r = pp[:, None, :]  # (kl, 1, cd)
        PQ2 = np.sum((PP[ij][None, :, None, :] - PP[:, None, :, :]) ** 2, axis=-1)
        pref = 2.0 * np.pi ** 2.5 / (q * r * np.sqrt(q + r))
        val = pref * KK[ij][None, :, None] * KK[:, None, :] * boys0(q * r / (q + r) * PQ2)
        eri[ij] = np.sum(val, axis=(1, 2))
    eri = eri.reshape(nat, nat, nat, nat)
    enuc = 0.0
    for i in range(nat):
        for j in range(i + 1, nat):
            enuc += Z[i] * Z[j] / float(np.linalg.norm(coords[i] - coords[j]))
    return Integrals(S=S, T=T, V=V, eri=eri, enuc=enuc)


def rhf(ints: Integrals, nocc: int, max_iter: int = 200, tol: float = 1e-11):
    """Closed-shell Roothaan-Hall SCF with DIIS.  Returns (energy, mo_coeff, mo_energy, converged)."""
    h = ints.T + ints.V
    S = ints.S
    e_s, U = np.linalg.eigh(S)
    X = U @ np.diag(e_s ** -0.5) @ U.T
    eri = ints.eri

    def fock(D):
        J = np.einsum("pqrs,rs->pq", eri, D)
        Kx = np.einsum("prqs,rs->pq", eri, D)
        return h + 2.0 * J - Kx

    def solve(F):
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        return e, X @ Cp

    e, C = solve(h)
    D = C[:, :nocc] @ C[:, :nocc].T
    E_old = 0.0
    focks, errs = [], []
    converged = False
    for _ in range(max_iter):
        F = fock(D)
        E = float(np.sum(D * (h + F)) + ints.enuc)
        err = F @ D @ S - S @ D @ F
        focks.append(F)
        errs.append(err)
        if len(focks) > 8:
            focks.pop(0)
            errs.pop(0)
        if abs(E - E_old) < tol and np.max(np.abs(err)) < 1e-8:
            converged = True
            break
        E_old = E
        nd = len(focks)
        Bm = -np.ones((nd + 1, nd + 1))
        Bm[nd, nd] = 0.0
        for i in range(nd):
            for j in range(nd):
                Bm[i, j] = np.sum(errs[i] * errs[j])
        rhs = np.zeros(nd + 1)
        rhs[nd] = -1.0
        coef = np.linalg.solve(Bm, rhs)[:nd]
        e, C = solve(sum(ci * Fi for ci, Fi in zip(coef, focks)))
        D = C[:, :nocc] @ C[:, :nocc].T
    # deterministic column signs: largest-magnitude component positive
    idx = np.argmax(np.abs(C), axis=0)
    C = C * np.sign(C[idx, np.arange(C.shape[1])])[None, :]
    return E, C, e, converged


def ao_cholesky(eri: np.ndarray, max_error: float = 1e-5) -> np.ndarray:
    """Pivoted Cholesky of the (nao^2 x nao^2) ERI matrix, as trot/staging.py:234-346 runs it."""
    nao = eri.shape[0]
    M = eri.reshape(nao * nao, nao * nao)
    diag = np.diag(M).copy()
    nmax = nao * nao
    L = np.zeros((nmax + 1, nao * nao))
    nu = int(np.argmax(diag))
    delta_max = diag[nu]
    L[0] = M[:, nu] / np.sqrt(delta_max)
    approx = np.zeros(nao * nao)
    n = 0
    while abs(delta_max) > max_error and n < nmax:
        approx += L[n] * L[n]
        delta = diag - approx
        nu = int(np.argmax(np.abs(delta)))
        delta_max = abs(delta[nu])
        R = L[: n + 1, nu] @ L[: n + 1, :]
        L[n + 1] = (M[:, nu] - R) / np.sqrt(delta_max)
        n += 1
    return L[:n].reshape(n, nao, nao)


@dataclass
class StagedRhf:
    h0: float
    h1: np.ndarray  # (N, N) MO basis
    chol: np.ndarray  # (G, N, N) MO basis
    trial: np.ndarray  # (N, nocc)
    nocc: int
    e_rhf: float
    mo_coeff: np.ndarray
    ao: Integrals


def stage_rhf(coords: np.ndarray, nelec: int, chol_cut: float = 1e-5) -> StagedRhf:
    """h0 / h1 / chol / trial in the RHF MO basis (trot/staging.py:882-975, 1110-1116, 1153-1163)."""
    if nelec % 2:
        raise ValueError("closed shell only")
    nocc = nelec // 2
    ints = integrals(coords)
    e_rhf, C, _, ok = rhf(ints, nocc)
    if not ok:
        raise RuntimeError("SCF did not converge")
    h1 = C.T @ (ints.T + ints.V) @ C
    L = ao_cholesky(ints.eri, chol_cut)
    chol = np.einsum("ip,gij,jq->gpq", C, L, C, optimize=True)
    q, r = np.linalg.qr(C.T @ ints.S @ C)
    q = q * np.sign(np.diag(r))[None, :]
    return StagedRhf(h0=ints.enuc, h1=h1, chol=chol, trial=np.ascontiguousarray(q[:, :nocc]), nocc=nocc,
                     e_rhf=e_rhf, mo_coeff=C, ao=ints)
