"""Seeded small AFQMC problems shared by the CPU and GPU tests (numpy only)."""

from __future__ import annotations

import numpy as np


def problem(norb: int, nocc: int, nchol: int, nw: int, seed: int, nocc_b: int | None = None, noise: float = 0.05,
            chol_scale: float = 0.1, h0: float = 0.37):
    """Synthetic restricted-basis Hamiltonian, real trial, complex walkers near the trial."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((norb, norb)) / np.sqrt(norb)
    h1 = np.triu(a) + np.triu(a, 1).T
    b = rng.standard_normal((nchol, norb, norb))
    chol = 0.5 * (b + np.swapaxes(b, 1, 2)) * chol_scale * np.exp(-np.arange(nchol) / max(nchol, 1))[:, None, None]
    evecs = np.linalg.eigh(h1)[1]
    ca = np.ascontiguousarray(evecs[:, :nocc])
    wa = ca[None] + noise * (rng.standard_normal((nw, norb, nocc)) + 1j * rng.standard_normal((nw, norb, nocc)))
    out = dict(h0=h0, h1=h1, chol=chol, ca=ca, wa=wa)
    if nocc_b is not None:
        cb = np.ascontiguousarray(evecs[:, :nocc_b])
        wb = cb[None] + noise * (rng.standard_normal((nw, norb, nocc_b)) + 1j * rng.standard_normal((nw, norb, nocc_b)))
        out.update(cb=cb, wb=wb)
    out["fields"] = [rng.standard_normal((nw, nchol)) for _ in range(8)]
    return out


def rel(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / den


def rel_each(a, b, floor: float = 1e-6) -> float:
    """Largest elementwise relative error |a_i - b_i| / max(|b_i|, floor max|b|).

    Every entry must be right to the tolerance, not only the large ones: a small weight or overlap that is
    wrong by orders of magnitude fails here and would pass rel().  Entries below floor x the largest are held
    to that floor (masked weights are exact zeros on both sides).
    """
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        raise ValueError(f"shapes differ: {a.shape} vs {b.shape}")
    if a.size == 0:
        return 0.0
    big = float(np.max(np.abs(b)))
    den = np.maximum(np.abs(b), max(floor * big, 1e-300))
    return float(np.max(np.abs(a - b) / den))
