"""Seeded synthetic AFQMC inputs: Hamiltonians, trials, walkers and auxiliary fields.

The reference draws its auxiliary fields from JAX's threefry stream
(trot/prop/afqmc.py:91-93), which is not reproducible without JAX.  Every
input here is instead drawn from ``numpy.random.default_rng`` (PCG64) in a
fixed order, so the GPU path and the CPU oracle can be fed identical data:

    h1        symmetric, h1_ij = h1_ji = a_ij / sqrt(N) for i <= j, a ~ N(0, 1)
    L_gamma   (A_gamma + A_gamma^T) / 2 * 0.1 * exp(-gamma / tau), A_gamma ~ N(0, 1),
              drawn for gamma = 0 .. G-1 after h1
    h0        0
    trial     lowest n eigenvectors of h1 (real)
    walkers   trial + 1e-2 (xi + i eta), xi, eta ~ N(0, 1), drawn after the L_gamma
    fields    N(0, 1) of shape (W, G), one generator per (seed, step)

``device`` arguments put the result on a torch device in chunks, so the
largest workload (N=512, G=2560: 5.4 GB of Cholesky vectors) never needs a
second host copy.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SyntheticSystem:
    h0: float
    h1: np.ndarray  # (N, N) float64
    chol: object  # (G, N, N) float64, numpy array or torch tensor
    trial: np.ndarray  # (N, n) float64 (RHF) ; for UHF see trial_b
    walkers: object  # (W, N, n) complex128, numpy or torch
    seed: int
    trial_b: np.ndarray | None = None
    walkers_b: object | None = None

    @property
    def norb(self) -> int:
        return self.h1.shape[0]

    @property
    def nchol(self) -> int:
        return int(self.chol.shape[0])


def _sym_h1(rng: np.random.Generator, norb: int) -> np.ndarray:
    a = rng.standard_normal((norb, norb)) / np.sqrt(norb)
    up = np.triu(a)
    return up + np.triu(a, 1).T


def _chol_chunks(rng: np.random.Generator, norb: int, nchol: int, tau: float, chunk: int = 64):
    for g0 in range(0, nchol, chunk):
        g1 = min(nchol, g0 + chunk)
        a = rng.standard_normal((g1 - g0, norb, norb))
        scale = 0.1 * np.exp(-np.arange(g0, g1) / tau)
        yield g0, g1, 0.5 * (a + np.swapaxes(a, 1, 2)) * scale[:, None, None]


def make_synthetic(norb: int, nocc: int, nchol: int, n_walkers: int, seed: int, tau: float,
                   noise: float = 1e-2, device=None, nocc_b: int | None = None,
                   chol_scale: float = 1.0) -> SyntheticSystem:
    """Synthetic restricted-basis Hamiltonian, RHF (or UHF with ``nocc_b``) trial and walkers.

    ``chol_scale`` multiplies every L_gamma after it is drawn (the draws are
    unchanged).  With the recipe's 0.1 amplitude and no 1/sqrt(N) the two-body
    part dominates so strongly at N >= 128 that h1_eff = h1 - 1/2 sum L^2 - ...
    has eigenvalues of -30 .. -50 against +-2 for h1, and every walker's
    importance leaves [floor, cap] within ~10 steps (in the CPU oracle as on
    the GPU); ``chol_scale = 1/sqrt(N)`` gives a system whose walkers survive,
    used for energy comparisons.  Flop counts do not depend on it.
    """
    rng = np.random.default_rng(seed)
    h1 = _sym_h1(rng, norb)
    if device is None:
        chol = np.empty((nchol, norb, norb))
        for g0, g1, blk in _chol_chunks(rng, norb, nchol, tau):
            chol[g0:g1] = blk * chol_scale
    else:
        import torch

        chol = torch.empty((nchol, norb, norb), dtype=torch.float64, device=device)
        for g0, g1, blk in _chol_chunks(rng, norb, nchol, tau):
            chol[g0:g1].copy_(torch.from_numpy(blk * chol_scale))
    evals, evecs = np.linalg.eigh(h1)
    trial = np.ascontiguousarray(evecs[:, :nocc])
    walkers = trial[None] + noise * (rng.standard_normal((n_walkers, norb, nocc))
                                     + 1j * rng.standard_normal((n_walkers, norb, nocc)))
    trial_b = walkers_b = None
    if nocc_b is not None:
        trial_b = np.ascontiguousarray(evecs[:, :nocc_b])
        walkers_b = trial_b[None] + noise * (rng.standard_normal((n_walkers, norb, nocc_b))
                                             + 1j * rng.standard_normal((n_walkers, norb, nocc_b)))
    if device is not None:
        import torch

        walkers = torch.from_numpy(walkers).to(device)
        if walkers_b is not None:
            walkers_b = torch.from_numpy(walkers_b).to(device)
    return SyntheticSystem(h0=0.0, h1=h1, chol=chol, trial=trial, walkers=walkers, seed=seed,
                           trial_b=trial_b, walkers_b=walkers_b)


def make_synthetic_cisd(norb: int, nocc: int, nchol: int, n_walkers: int, seed: int, tau: float, device,
                        ci1_scale: float = 0.02, ci2_scale: float = 0.005, noise: float = 1e-2):
    """A CISD workload: the recipe's Hamiltonian (every L_gamma scaled by 1/sqrt(N) so that walkers
    survive) rotated to the eigenbasis of h1, where the reference determinant occupies the first nocc
    orbitals (the basis the CISD trial assumes); random real amplitudes ci1 (nocc, nvir) and ci2 with
    pair symmetry ci2[i,a,j,b] = ci2[j,b,i,a]; walkers near [I; 0].  Returns (h0, h1, chol, ci1, ci2,
    walkers) with h1, chol and walkers as GPU tensors."""
    import torch

    syn = make_synthetic(norb, nocc, nchol, 1, seed, tau, device=device, chol_scale=1.0 / np.sqrt(norb))
    evals, evecs = np.linalg.eigh(syn.h1)
    c = torch.from_numpy(evecs).to(device)
    chol = torch.matmul(torch.matmul(c.T, syn.chol), c)  # C^T L_g C for every g
    chol = 0.5 * (chol + chol.transpose(1, 2))
    h1 = torch.from_numpy(np.diag(evals)).to(device)
    rng = np.random.default_rng(seed + 1000)
    nvir = norb - nocc
    ci1 = ci1_scale * rng.standard_normal((nocc, nvir))
    ci2 = ci2_scale * rng.standard_normal((nocc, nvir, nocc, nvir))
    ci2 = 0.5 * (ci2 + ci2.transpose(2, 3, 0, 1))
    w = np.zeros((n_walkers, norb, nocc), complex)
    w[:, :nocc] = np.eye(nocc)
    w = w + noise * (rng.standard_normal(w.shape) + 1j * rng.standard_normal(w.shape))
    return syn.h0, h1, chol.contiguous(), ci1, ci2, torch.from_numpy(w).to(device)


def fields(seed: int, step: int, n_walkers: int, nchol: int) -> np.ndarray:
    """Pre-drawn N(0,1) auxiliary fields for one step, (W, G) float64."""
    return np.random.default_rng([seed, 7919, step]).standard_normal((n_walkers, nchol))


def sr_zeta(seed: int, block: int) -> float:
    """Pre-drawn U(0,1) comb offset for the stochastic reconfiguration of one block."""
    return float(np.random.default_rng([seed, 104729, block]).uniform())


# Workloads named by their sizes (N orbitals / n occupied / G Cholesky vectors / W walkers).
WORKLOADS = {
    "n256_o32_g1024_w1024": dict(norb=256, nocc=32, nchol=1024, n_walkers=1024, seed=1, tau=256.0, dt=0.01),
    "n512_o64_g2560_w2048": dict(norb=512, nocc=64, nchol=2560, n_walkers=2048, seed=2, tau=512.0, dt=0.005),
    "n128_o16_g640_w512": dict(norb=128, nocc=16, nchol=640, n_walkers=512, seed=3, tau=256.0, dt=0.005),
    # CISD trial (make_synthetic_cisd): a mid-size molecule in a double-zeta basis
    "cisd_n128_o16_g512_w1024": dict(norb=128, nocc=16, nchol=512, n_walkers=1024, seed=5, tau=256.0, dt=0.005),
}


def fragment_sizes(n_frag: int = 32, seed: int = 4, lo: int = 40, hi: int = 200):
    """Uneven fragment sizes: N ~ U[lo, hi], n = round(N/5), G = 4N."""
    rng = np.random.default_rng(seed)
    norbs = rng.integers(lo, hi + 1, size=n_frag)
    return [(int(nb), int(round(nb / 5)), int(4 * nb)) for nb in norbs]
