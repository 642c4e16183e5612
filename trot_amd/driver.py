"""Equilibration + sampling driver (trot/driver.py:168-406) over GPU blocks.

Blocks run back to back on the GPU; block scalars are pulled to the host once
per print chunk, as the reference does after each jitted chunk of blocks.
"""

from __future__ import annotations

import time
from functools import partial
from typing import Any, NamedTuple

import numpy as np

from .blocks import block as default_block
from .stats import blocking_analysis_ratio, reject_outliers

print = partial(print, flush=True)


class QmcResult(NamedTuple):
    """trot/driver.py:33-40 (observables are not part of this path: always empty)."""

    mean_energy: float
    stderr_energy: float | None
    block_energies: np.ndarray
    block_weights: np.ndarray
    block_observables: dict
    observable_means: dict
    observable_stderrs: dict


def run_qmc(*, sys, params, ham_data, trial_data, meas_ops, trial_ops, prop_ops, block_fn=default_block,
            state=None, meas_ctx=None, prop_ctx=None, target_error: float | None = None, mesh=None,
            observable_names: tuple = (), verbose: bool = True) -> QmcResult:
    """trot/driver.py:168-406: n_eql_blocks equilibration blocks, then n_blocks sampling blocks."""
    if mesh is not None:
        raise NotImplementedError("trot_amd runs on one GPU")
    for name in observable_names:
        meas_ops.require_observable(name)
    if prop_ctx is None:
        prop_ctx = prop_ops.build_prop_ctx(ham_data, trial_ops.get_rdm1(trial_data), params)
    if meas_ctx is None:
        meas_ctx = meas_ops.build_meas_ctx(ham_data, trial_data)
    if state is None:
        state = prop_ops.init_prop_state(sys=sys, ham_data=ham_data, trial_ops=trial_ops, trial_data=trial_data,
                                         meas_ops=meas_ops, params=params)
    log = print if verbose else (lambda *a, **k: None)

    def run_blocks(st, n):
        es, ws = [], []
        for _ in range(n):
            st, obs = block_fn(st, sys=sys, params=params, ham_data=ham_data, trial_data=trial_data,
                               trial_ops=trial_ops, meas_ops=meas_ops, meas_ctx=meas_ctx, prop_ops=prop_ops,
                               prop_ctx=prop_ctx)
            es.append(obs.scalars["energy"])
            ws.append(obs.scalars["weight"])
        import torch

        e = torch.stack(es).cpu().numpy() if es else np.zeros(0)
        w = torch.stack(ws).cpu().numpy() if ws else np.zeros(0)
        return st, e, w

    t0 = time.perf_counter()
    block_e_eq = [float(state.e_estimate.item())]
    block_w_eq = [float(state.weights.sum().item())]
    log("\nEquilibration:\n")
    chunk = max(1, params.n_eql_blocks // 5)
    for start in range(0, params.n_eql_blocks, chunk):
        n = min(chunk, params.n_eql_blocks - start)
        state, e, w = run_blocks(state, n)
        block_e_eq.extend(e.tolist())
        block_w_eq.extend(w.tolist())
        log(f"[eql {start + n:4d}/{params.n_eql_blocks}]  {float(np.mean(e * w) / np.mean(w)):14.10f}  "
            f"{float(np.mean(w)):12.6e}  {int(state.node_encounters.item()):10d}  {time.perf_counter() - t0:8.1f}")

    log("\nSampling:\n")
    target_error = 0.0 if target_error is None else target_error
    block_e_s: list[float] = []
    block_w_s: list[float] = []
    chunk = max(1, params.n_blocks // 10)
    t_mark = time.perf_counter()
    for start in range(0, params.n_blocks, chunk):
        n = min(chunk, params.n_blocks - start)
        state, e, w = run_blocks(state, n)
        block_e_s.extend(e.tolist())
        block_w_s.extend(w.tolist())
        stats = blocking_analysis_ratio(np.asarray(block_e_s), np.asarray(block_w_s), print_q=False)
        se = stats["se_star"]
        dt_blk = (time.perf_counter() - t_mark) / n
        t_mark = time.perf_counter()
        log(f"[blk {start + n:4d}/{params.n_blocks}]  {stats['mu']:14.10f}  "
            f"{(f'{se:10.3e}' if se is not None else ' ' * 10)}  {float(np.mean(e * w) / np.mean(w)):16.10f}  "
            f"{float(np.mean(w)):12.6e}  {int(state.node_encounters.item()):10d}  {dt_blk:9.3f}  "
            f"{time.perf_counter() - t0:8.1f}")
        if se is not None and target_error > 0.0 and se <= target_error:
            log(f"\nTarget error {target_error:.3e} reached at block {start + n}.")
            break
    data = np.column_stack((block_e_s, block_w_s))
    clean, keep = reject_outliers(data, obs=0)
    log(f"\nRejected {data.shape[0] - clean.shape[0]} outlier blocks.")
    stats = blocking_analysis_ratio(clean[:, 0], clean[:, 1], print_q=verbose)
    return QmcResult(
        mean_energy=stats["mu"],
        stderr_energy=stats["se_star"],
        block_energies=np.concatenate([block_e_eq, clean[:, 0]]),
        block_weights=np.concatenate([block_w_eq, clean[:, 1]]),
        block_observables={},
        observable_means={},
        observable_stderrs={},
    )
