"""Brute-force Fock-space evaluation for tiny systems (N <= 6 orbitals).

TEST INFRASTRUCTURE ONLY (see ``oracle/afqmc_oracle.py``).  This is an
independent check of the half-rotated force bias and local energy formulas
(trot/meas/rhf.py:119-172, trot/meas/uhf.py:43-57, 205-231): it never uses
Green's functions.  It expands Slater determinants into occupation-number
vectors (coefficient of a string = minor of the orbital matrix), builds the
second-quantised Hamiltonian

    H = h0 + sum_pq h1_pq E_pq + 1/2 sum_pqrs (pq|rs) (E_pq E_rs - delta_qr E_ps),
    (pq|rs) = sum_g L_g,pq L_g,rs,   E_pq = sum_sigma a+_{p sigma} a_{q sigma},

as an explicit matrix, and evaluates <T|O|w> / <T|w> directly.
"""

from __future__ import annotations

from itertools import combinations

import numpy as np


def _strings(norb: int, nel: int) -> list[tuple[int, ...]]:
    return list(combinations(range(norb), nel))


def _one_spin_epq(norb: int, nel: int) -> np.ndarray:
    """Matrices of a+_p a_q on the strings of one spin, shape (N, N, S, S)."""
    strs = _strings(norb, nel)
    index = {s: i for i, s in enumerate(strs)}
    ns = len(strs)
    out = np.zeros((norb, norb, ns, ns))
    for j, s in enumerate(strs):
        occ = set(s)
        for q in s:
            sign_q = (-1) ** sum(1 for o in s if o < q)
            rest = sorted(occ - {q})
            for p in range(norb):
                if p in rest:
                    continue
                sign_p = (-1) ** sum(1 for o in rest if o < p)
                t = tuple(sorted(rest + [p]))
                out[p, q, index[t], j] += sign_q * sign_p
    return out


def _amplitudes(orbs: np.ndarray, nel: int) -> np.ndarray:
    """Coefficients of prod_k (sum_p orbs[p,k] a+_p)|0> on ascending strings: minors."""
    norb = orbs.shape[0]
    return np.array([np.linalg.det(orbs[list(s), :]) for s in _strings(norb, nel)], dtype=complex)


class FockSpace:
    """Determinant space with fixed (n_up, n_dn) on N spatial orbitals."""

    def __init__(self, norb: int, nelec: tuple[int, int]):
        self.norb = norb
        self.nelec = nelec
        eu = _one_spin_epq(norb, nelec[0])
        ed = _one_spin_epq(norb, nelec[1])
        nu, nd = eu.shape[-1], ed.shape[-1]
        iu, idn = np.eye(nu), np.eye(nd)
        # E_pq (both spins) on the product space, ordered (up string, down string)
        self.epq = np.einsum("pqab,cd->pqacbd", eu, idn).reshape(norb, norb, nu * nd, nu * nd)
        self.epq = self.epq + np.einsum("ab,pqcd->pqacbd", iu, ed).reshape(norb, norb, nu * nd, nu * nd)

    def state(self, w_up: np.ndarray, w_dn: np.ndarray) -> np.ndarray:
        """|w_up> x |w_dn> as a vector (up string major)."""
        return np.kron(_amplitudes(w_up, self.nelec[0]), _amplitudes(w_dn, self.nelec[1]))

    def one_body(self, m: np.ndarray) -> np.ndarray:
        return np.einsum("pq,pqab->ab", m, self.epq)

    def hamiltonian(self, h0: float, h1: np.ndarray, chol: np.ndarray) -> np.ndarray:
        dim = self.epq.shape[-1]
        h = h0 * np.eye(dim) + self.one_body(h1)
        eri_corr = 0.5 * np.einsum("gpq,gqs->ps", chol, chol)
        h = h - self.one_body(eri_corr)
        for lg in chol:
            op = self.one_body(lg)
            h = h + 0.5 * op @ op
        return h

    def mixed(self, op: np.ndarray, bra: np.ndarray, ket: np.ndarray) -> complex:
        return complex(bra.conj() @ op @ ket / (bra.conj() @ ket))


def local_energy(h0, h1, chol, trial_up, trial_dn, w_up, w_dn) -> complex:
    """<T|H|w>/<T|w> by explicit matrices."""
    fs = FockSpace(h1.shape[0], (trial_up.shape[1], trial_dn.shape[1]))
    bra = fs.state(trial_up, trial_dn)
    ket = fs.state(w_up, w_dn)
    return fs.mixed(fs.hamiltonian(h0, h1, chol), bra, ket)


def force_bias(chol, trial_up, trial_dn, w_up, w_dn) -> np.ndarray:
    """<T|L_g|w>/<T|w> for every Cholesky vector, L_g = sum_pq L_g,pq E_pq."""
    fs = FockSpace(chol.shape[1], (trial_up.shape[1], trial_dn.shape[1]))
    bra = fs.state(trial_up, trial_dn)
    ket = fs.state(w_up, w_dn)
    return np.array([fs.mixed(fs.one_body(lg), bra, ket) for lg in chol])


def overlap(trial_up, trial_dn, w_up, w_dn) -> complex:
    fs = FockSpace(trial_up.shape[0], (trial_up.shape[1], trial_dn.shape[1]))
    return complex(fs.state(trial_up, trial_dn).conj() @ fs.state(w_up, w_dn))
