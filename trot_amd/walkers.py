"""Walker utilities (trot/walkers.py): init, batching helpers, QR, stochastic reconfiguration."""

from __future__ import annotations

from typing import Callable

import numpy as np
import torch

from . import _lib
from .core import System


def _natorbs(dm: np.ndarray, n_occ: int) -> np.ndarray:
    """trot/walkers.py:14-17 (host, N x N)."""
    dm = 0.5 * (dm + dm.conj().T)
    vecs = np.linalg.eigh(dm)[1][:, ::-1]
    return vecs[:, :n_occ]


def init_walkers(sys: System, rdm1, n_walkers: int):
    """trot/walkers.py:20-78 for restricted / unrestricted walkers (complex128 on the GPU)."""
    dm = rdm1.detach().cpu().numpy() if isinstance(rdm1, torch.Tensor) else np.asarray(rdm1)
    if dm.ndim != 3 or dm.shape[0] != 2:
        raise ValueError("Expected rdm1 with shape (2, norb, norb).")
    kind = sys.walker_kind
    if kind == "restricted":
        w0 = _natorbs(dm[0] + dm[1], sys.nup) + 0.0j
        return torch.from_numpy(np.ascontiguousarray(np.broadcast_to(w0, (n_walkers, *w0.shape)))).to("cuda")
    if kind == "unrestricted":
        wu = _natorbs(dm[0], sys.nup) + 0.0j
        wd = _natorbs(dm[1], sys.ndn) + 0.0j
        return (torch.from_numpy(np.ascontiguousarray(np.broadcast_to(wu, (n_walkers, *wu.shape)))).to("cuda"),
                torch.from_numpy(np.ascontiguousarray(np.broadcast_to(wd, (n_walkers, *wd.shape)))).to("cuda"))
    if kind == "generalized":
        raise NotImplementedError("generalized walkers are not part of the trot_amd GPU path")
    raise ValueError(f"unknown walker_kind: {kind}")


def as_device_walkers(w):
    """complex128, contiguous, on the GPU."""
    def one(x):
        if isinstance(x, torch.Tensor):
            return x.to(device="cuda", dtype=torch.complex128).contiguous()
        return torch.as_tensor(np.asarray(x), dtype=torch.complex128, device="cuda").contiguous()

    if isinstance(w, tuple):
        return (one(w[0]), one(w[1]))
    return one(w)


def is_unrestricted(w) -> bool:
    return isinstance(w, tuple) and len(w) == 2


def n_walkers(w) -> int:
    return int((w[0] if isinstance(w, tuple) else w).shape[0])


def take_walkers(w, idx):
    """trot/walkers.py:328-329."""
    idx = torch.as_tensor(idx, device="cuda", dtype=torch.long)
    if isinstance(w, tuple):
        return (w[0][idx].contiguous(), w[1][idx].contiguous())
    return w[idx].contiguous()


def vmap_chunked(fn: Callable, n_chunks: int, *, in_axes=0):
    """trot/walkers.py:94-134.  trot_amd kernels are batched already: the call goes straight through."""

    def wrapped(*args, **kwargs):
        return fn(*args, **kwargs)

    return wrapped


def orthonormalize(w, walker_kind: str):
    """trot/walkers.py:151-179 (norms discarded): Q with real positive diag(R)."""
    kind = walker_kind.lower()
    if kind == "unrestricted":
        return (_lib.orthonormalize(w[0]), _lib.orthonormalize(w[1]))
    if kind == "restricted":
        return _lib.orthonormalize(w)
    raise ValueError(f"unknown walker_kind: {walker_kind}")


def sr_indices(weights: torch.Tensor, zeta: float) -> tuple[torch.Tensor, torch.Tensor]:
    """trot/walkers.py:249-254 on the GPU: (idx (W,) int32, new weights (W,))."""
    return _lib.sr_comb(weights, zeta)


def gather(x, idx: torch.Tensor):
    """Row gather x[idx] on the GPU for tensors or pairs of tensors (None passes through)."""
    if x is None:
        return None
    if isinstance(x, tuple):
        return tuple(gather(t, idx) for t in x)
    return _lib.gather_rows(x, idx)


def stochastic_reconfiguration(w, weights: torch.Tensor, zeta: float, walker_kind: str, *, data_sharding=None):
    """trot/walkers.py:257-293: comb over cumsum(|w|); returns (walkers[idx], full(sum|w|/W))."""
    if data_sharding is not None:
        raise NotImplementedError("trot_amd runs on one GPU")
    idx, w_new = sr_indices(weights, zeta)
    return gather(w, idx), w_new


def no_sr(w, weights, zeta, walker_kind: str):
    """trot/walkers.py:240-246."""
    return w, weights
