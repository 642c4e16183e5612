"""Pins of the differentiable restatement (oracle/autodiff.py), which stands in for the reference's jax.grad
of the block energy (README.md:11; trot/prop/afqmc.py:79-144, trot/prop/blocks.py:236-257):

* its forward block energy equals the numpy oracle's ``block`` (oracle/afqmc_oracle.py) at 1e-11 relative on a
  small restricted-walker system with an RHF trial, with no importance masked, no weight capped and no energy
  clamped;
* its autograd gradient dE/dh1, symmetrised, against central differences of the numpy oracle's block energy
  along random symmetric directions D (1e-7 relative), and along single (i, j) + (j, i) entries.
"""

from __future__ import annotations

import numpy as np
import torch

from oracle import afqmc_oracle as O
from oracle import autodiff


def _system(seed=5, N=10, n=3, G=12, W=6, steps=8):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((N, N)) / np.sqrt(N)
    h1 = np.triu(a) + np.triu(a, 1).T
    b = rng.standard_normal((G, N, N))
    chol = 0.15 * (b + np.swapaxes(b, 1, 2)) * np.exp(-np.arange(G) / G)[:, None, None]
    c = np.linalg.eigh(h1)[1][:, :n]
    w0 = c[None] + 0.05 * (rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n)))
    fields = [rng.standard_normal((W, G)) for _ in range(steps)]
    return dict(h0=0.3, h1=h1, chol=chol, c=c, w0=w0, fields=fields, W=W, steps=steps, dt=0.01)


def _oracle_energy(s, h1):
    params = O.QmcParams(dt=s["dt"], n_walkers=s["W"], n_prop_steps=s["steps"])
    ctx = O.build_prop_ctx(s["h0"], h1, s["chol"], O.rhf_rdm1(s["c"]), s["dt"])
    ops = O.make_rhf_ops(s["h0"], h1, s["chol"], s["c"])
    st = O.init_prop_state(s["w0"], ops, s["W"])
    _, obs = O.block(st, s["fields"], 0.5, ctx, ops, params)
    return obs["energy"]


def _autograd(s):
    h1t = torch.tensor(s["h1"], dtype=torch.float64, requires_grad=True)
    e, info = autodiff.block_energy(s["h0"], h1t, s["chol"], s["c"], s["w0"], s["fields"], s["dt"])
    e.backward()
    g = h1t.grad.numpy()
    return float(e.item()), 0.5 * (g + g.T), info


def test_forward_matches_oracle_block():
    s = _system()
    e, _, info = _autograd(s)
    assert info == {"masked_importance": 0, "capped_weight": 0, "clamped_energy": 0}
    e_ref = _oracle_energy(s, s["h1"])
    assert abs(e - e_ref) <= 1e-11 * abs(e_ref)  # measured 6e-13: torch and numpy LAPACK paths


def test_gradient_matches_central_differences_of_the_oracle():
    s = _system()
    _, grad, _ = _autograd(s)
    rng = np.random.default_rng(11)
    h = 1e-5
    N = s["h1"].shape[0]
    dirs = []
    for _ in range(2):
        d = rng.standard_normal((N, N))
        dirs.append(0.1 * (d + d.T))
    for (i, j) in [(0, 0), (2, 7)]:
        d = np.zeros((N, N))
        d[i, j] += 1.0
        if i != j:
            d[j, i] += 1.0
        dirs.append(d)
    for D in dirs:
        fd = (_oracle_energy(s, s["h1"] + h * D) - _oracle_energy(s, s["h1"] - h * D)) / (2 * h)
        ad = float(np.sum(grad * D))
        assert abs(ad - fd) <= 1e-7 * max(abs(fd), 1e-3 * np.abs(grad).max()), (ad, fd)
