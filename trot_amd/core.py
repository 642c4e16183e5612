"""Core protocols and static system data, mirroring trot/core and trot/ham/chol.py.

Arrays are torch tensors on the GPU.  Measurement kernels are batched over a
leading walker axis (the reference maps per-walker kernels with
``vmap_chunked``, trot/walkers.py:94-134); calling one with a single walker
(no leading axis) is also accepted.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable, Literal, Mapping, NamedTuple

import numpy as np
import torch

HamBasis = Literal["restricted", "generalized"]
WalkerKind = Literal["restricted", "unrestricted", "generalized"]

# usual kernel names (trot/core/ops.py:116-120)
k_energy = "energy"
k_force_bias = "force_bias"
o_rdm1 = "rdm1"
o_density_corr = "density_corr"


@dataclass(frozen=True)
class System:
    """trot/core/system.py:9-35."""

    norb: int
    nelec: tuple[int, int]
    walker_kind: str

    def __post_init__(self):
        object.__setattr__(self, "walker_kind", self.walker_kind.lower())

    @property
    def nup(self) -> int:
        return self.nelec[0]

    @property
    def ndn(self) -> int:
        return self.nelec[1]

    @property
    def ne(self) -> int:
        return self.nelec[0] + self.nelec[1]


def _as_gpu(x, dtype=torch.float64, device=None) -> torch.Tensor:
    dev = device if device is not None else torch.device("cuda")
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=dev).contiguous()


@dataclass(frozen=True)
class HamChol:
    """Cholesky Hamiltonian (trot/ham/chol.py:13-58): h0, h1 (N,N), chol (G,N,N), real.

    ``chol`` may be dropped (``None``) after the propagation and measurement
    contexts are built, as the reference's runtime layout does
    (trot/runtime_layout.py:89-118).
    """

    h0: Any
    h1: Any
    chol: Any
    basis: str = "restricted"
    nchol: int | None = None

    def __post_init__(self):
        if self.basis not in ("restricted", "generalized"):
            raise ValueError(f"unknown basis: {self.basis}")
        shape = getattr(self.chol, "shape", None)
        if shape is None:
            return
        if self.nchol is None:
            object.__setattr__(self, "nchol", int(shape[0]))
        elif int(shape[0]) not in (0, int(self.nchol)):
            raise ValueError(f"nchol={self.nchol} is inconsistent with chol.shape[0]={shape[0]}")

    @property
    def norb(self) -> int:
        return int(self.h1.shape[0])

    def to_device(self, device=None) -> "HamChol":
        """Copy h1 / chol to the GPU as contiguous float64 tensors."""
        chol = None if self.chol is None else _as_gpu(self.chol, device=device)
        return HamChol(h0=float(np.real(np.asarray(self.h0 if not isinstance(self.h0, torch.Tensor) else
                                                   self.h0.cpu()))),
                       h1=_as_gpu(self.h1, device=device), chol=chol, basis=self.basis, nchol=self.nchol)


def n_fields(ham: HamChol) -> int:
    """trot/ham/chol.py:61-64."""
    assert ham.nchol is not None
    return int(ham.nchol)


class TrialOps(NamedTuple):
    """trot/core/ops.py:49-68 (CPMC fast-update hooks are not part of this path)."""

    overlap: Callable
    get_rdm1: Callable
    calc_green: Callable | None = None
    calc_overlap_ratio: Callable | None = None
    update_green: Callable | None = None


@dataclass(frozen=True)
class MeasOps:
    """trot/core/ops.py:124-166.

    ``engine`` is the batched GPU implementation behind the kernels; the step
    uses it to reuse the half Green's function computed with the overlap.
    """

    overlap: Callable
    build_meas_ctx: Callable = lambda ham_data, trial_data: None
    kernels: Mapping[str, Callable] = field(default_factory=dict)
    observables: Mapping[str, Callable] = field(default_factory=dict)
    engine: Any = None

    def has_kernel(self, name: str) -> bool:
        return name in self.kernels

    def has_observable(self, name: str) -> bool:
        return name in self.observables

    def require_kernel(self, name: str) -> Callable:
        try:
            return self.kernels[name]
        except KeyError as e:
            avail = ", ".join(sorted(self.kernels.keys()))
            raise KeyError(f"missing required kernel '{name}'. available: [{avail}]") from e

    def require_observable(self, name: str) -> Callable:
        try:
            return self.observables[name]
        except KeyError as e:
            avail = ", ".join(sorted(self.observables.keys()))
            raise KeyError(f"missing requested observable '{name}'. available: [{avail}]") from e

    def available_kernels(self) -> tuple[str, ...]:
        return tuple(sorted(self.kernels.keys()))

    def available_observables(self) -> tuple[str, ...]:
        return tuple(sorted(self.observables.keys()))
