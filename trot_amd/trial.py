"""RHF / UHF trials (trot/trial/rhf.py, trot/trial/uhf.py) with GPU overlaps.

Trial orbitals must be real (the molecular RHF/UHF case the reference's
staging produces); complex trial orbitals are rejected with ValueError.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .core import System, TrialOps


def _real_orbitals(c, device=None) -> torch.Tensor:
    if isinstance(c, torch.Tensor):
        if c.is_complex():
            if torch.count_nonzero(c.imag).item():
                raise ValueError("trot_amd supports real trial orbitals only")
            c = c.real
        return c.to(device=device or "cuda", dtype=torch.float64).contiguous()
    c = np.asarray(c)
    if np.iscomplexobj(c):
        if np.any(c.imag != 0):
            raise ValueError("trot_amd supports real trial orbitals only")
        c = c.real
    return torch.as_tensor(c, dtype=torch.float64, device=device or "cuda").contiguous()


@dataclass(frozen=True)
class RhfTrial:
    """trot/trial/rhf.py:13-32: mo_coeff (norb, nocc)."""

    mo_coeff: torch.Tensor

    def __post_init__(self):
        object.__setattr__(self, "mo_coeff", _real_orbitals(self.mo_coeff))

    @property
    def norb(self) -> int:
        return int(self.mo_coeff.shape[0])

    @property
    def nocc(self) -> int:
        return int(self.mo_coeff.shape[1])


@dataclass(frozen=True)
class UhfTrial:
    """trot/trial/uhf.py:13-33: mo_coeff_a (norb, na), mo_coeff_b (norb, nb)."""

    mo_coeff_a: torch.Tensor
    mo_coeff_b: torch.Tensor

    def __post_init__(self):
        object.__setattr__(self, "mo_coeff_a", _real_orbitals(self.mo_coeff_a))
        object.__setattr__(self, "mo_coeff_b", _real_orbitals(self.mo_coeff_b))

    @property
    def norb(self) -> int:
        return int(self.mo_coeff_a.shape[0])

    @property
    def nocc(self) -> tuple[int, int]:
        return (int(self.mo_coeff_a.shape[1]), int(self.mo_coeff_b.shape[1]))


def _batched(w):
    """Accept one walker (N, n) or a batch (W, N, n); returns (batch, squeeze flag)."""
    if isinstance(w, tuple):
        wu, wd = w
        if wu.dim() == 2:
            return (wu.unsqueeze(0).contiguous(), wd.unsqueeze(0).contiguous()), True
        return (wu.contiguous(), wd.contiguous()), False
    if w.dim() == 2:
        return w.unsqueeze(0).contiguous(), True
    return w.contiguous(), False


def rhf_get_rdm1(trial_data: RhfTrial) -> torch.Tensor:
    """trot/trial/rhf.py:39-42."""
    c = trial_data.mo_coeff
    dm = c @ c.T
    return torch.stack([dm, dm], dim=0)


def uhf_get_rdm1(trial_data: UhfTrial) -> torch.Tensor:
    """trot/trial/uhf.py:40-45."""
    ca, cb = trial_data.mo_coeff_a, trial_data.mo_coeff_b
    return torch.stack([ca @ ca.T, cb @ cb.T], dim=0)


def _coeffs(trial_data):
    if isinstance(trial_data, UhfTrial):
        return trial_data.mo_coeff_a, trial_data.mo_coeff_b
    return trial_data.mo_coeff, trial_data.mo_coeff


def overlap_green(walkers, trial_data, want_green: bool = True, out=None):
    """Batched overlaps and transposed half Green's functions.

    Returns (overlap (W,), (GT_a, GT_b or None)).  Restricted walkers give
    det(C^T w)^2 (trot/trial/rhf.py:45-47); unrestricted give det_a det_b
    (trot/trial/rhf.py:50-54, trot/trial/uhf.py:54-58).  ``out`` (GT_a, GT_b or None): buffers that receive
    the Green's functions.
    """
    ca, cb = _coeffs(trial_data)
    oa, ob = out if out is not None else (None, None)
    if isinstance(walkers, tuple):
        det_a, gt_a = _lib.overlap_green(ca, walkers[0], want_green, oa)
        det_b, gt_b = _lib.overlap_green(cb, walkers[1], want_green, ob)
        return _lib.overlap_combine(det_a, det_b), (gt_a, gt_b)
    det_a, gt_a = _lib.overlap_green(ca, walkers, want_green, oa)
    return _lib.overlap_combine(det_a, None), (gt_a, None)


def overlap(walker, trial_data) -> torch.Tensor:
    """Overlap <T|w> for one walker or a batch (the TrialOps.overlap role)."""
    w, squeeze = _batched(walker)
    ov, _ = overlap_green(w, trial_data, want_green=False)
    return ov[0] if squeeze else ov


def make_rhf_trial_ops(sys: System) -> TrialOps:
    """trot/trial/rhf.py:66-87."""
    if sys.nup != sys.ndn:
        raise ValueError("RHF requires nelec[0] == nelec[1].")
    wk = sys.walker_kind
    if wk == "generalized":
        raise NotImplementedError("generalized walkers are not part of the trot_amd GPU path")
    if wk not in ("restricted", "unrestricted"):
        raise ValueError(f"unknown walker_kind: {sys.walker_kind}")
    return TrialOps(overlap=overlap, get_rdm1=rhf_get_rdm1)


def make_uhf_trial_ops(sys: System) -> TrialOps:
    """trot/trial/uhf.py:71-89 (unrestricted walkers)."""
    wk = sys.walker_kind
    if wk in ("restricted", "generalized"):
        raise NotImplementedError(f"UHF trial with {wk} walkers is not part of the trot_amd GPU path")
    if wk != "unrestricted":
        raise ValueError(f"unknown walker_kind: {sys.walker_kind}")
    return TrialOps(overlap=overlap, get_rdm1=uhf_get_rdm1)


def make_rhf_trial_data(data: dict, sys: System) -> RhfTrial:
    """trot/trial/rhf.py:90-93."""
    mo = np.asarray(data["mo"]) if not isinstance(data["mo"], torch.Tensor) else data["mo"]
    return RhfTrial(mo[:, : sys.nup])


def make_uhf_trial_data(data: dict, sys: System) -> UhfTrial:
    """trot/trial/uhf.py:92-105."""
    if "mo_a" in data and "mo_b" in data:
        mo_a, mo_b = data["mo_a"], data["mo_b"]
    elif "mo" in data:
        mo_a = mo_b = data["mo"]
    else:
        raise KeyError("Failed to find the trial coeff.")
    return UhfTrial(mo_a[:, : sys.nup], mo_b[:, : sys.ndn])
