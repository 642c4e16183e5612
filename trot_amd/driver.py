"""Equilibration + sampling driver (trot/driver.py:168-406) over GPU blocks.

Blocks run back to back on the GPU; block scalars are pulled to the host once
per print chunk, as the reference does after each jitted chunk of blocks.
"""

from __future__ import annotations

import time
from functools import partial
from typing import NamedTuple

import numpy as np

from .blocks import block as default_block
from .stats import blocking_analysis_ratio, jackknife_ratios, rebin_observable, reject_outliers

print = partial(print, flush=True)


class QmcResult(NamedTuple):
    """trot/driver.py:33-40.  block_observables: name -> (equilibration + kept sampling blocks, *shape) host
    arrays; observable_means / observable_stderrs: name -> weighted mean over the kept sampling blocks and
    its jackknife error at the energy's B* (NaN when that leaves fewer than 2 super-blocks)."""

    mean_energy: float
    stderr_energy: float | None
    block_energies: np.ndarray
    block_weights: np.ndarray
    block_observables: dict
    observable_means: dict
    observable_stderrs: dict


def make_run_blocks(*, block_fn, sys, params, trial_ops, meas_ops, prop_ops, observable_names: tuple = ()):
    """trot/driver.py:64-110: ``run_blocks(state, *, ham_data, trial_data, meas_ctx, prop_ctx, n_blocks)`` runs
    n_blocks blocks back to back on the GPU and returns (state, scalars, observables): scalars maps each
    block scalar ("energy", "weight") to an (n_blocks,) device tensor, observables is a tuple with one
    (n_blocks, ...) device tensor per name of ``observable_names`` (empty tuples / tensors for n_blocks = 0)."""
    import torch

    names = tuple(observable_names)
    extra = {"observable_names": names} if names else {}

    def run_blocks(state, *, ham_data, trial_data, meas_ctx, prop_ctx, n_blocks: int):
        scal, obs = {}, {name: [] for name in names}
        for _ in range(int(n_blocks)):
            state, o = block_fn(state, sys=sys, params=params, ham_data=ham_data, trial_data=trial_data,
                                trial_ops=trial_ops, meas_ops=meas_ops, meas_ctx=meas_ctx, prop_ops=prop_ops,
                                prop_ctx=prop_ctx, **extra)
            for k, v in o.scalars.items():
                scal.setdefault(k, []).append(v.reshape(()))
            for name in names:
                obs[name].append(o.observables[name])
        scalars = {k: torch.stack(v) for k, v in scal.items()}
        if not scalars:
            scalars = {"energy": torch.zeros(0, dtype=torch.float64), "weight": torch.zeros(0, dtype=torch.float64)}
        return state, scalars, tuple(torch.stack(obs[name]) if obs[name] else None for name in names)

    return run_blocks


def run_qmc_energy(*, sys, params, ham_data, trial_data, meas_ops, trial_ops, prop_ops, block_fn=default_block,
                   state=None, meas_ctx=None, prop_ctx=None, target_error: float | None = None, mesh=None,
                   verbose: bool = True):
    """trot/driver.py:771-803: run_qmc without observables; (mean energy, its error, block energies,
    block weights)."""
    out = run_qmc(sys=sys, params=params, ham_data=ham_data, trial_data=trial_data, meas_ops=meas_ops,
                  trial_ops=trial_ops, prop_ops=prop_ops, block_fn=block_fn, state=state, meas_ctx=meas_ctx,
                  prop_ctx=prop_ctx, target_error=target_error, mesh=mesh, observable_names=(), verbose=verbose)
    return out.mean_energy, out.stderr_energy, out.block_energies, out.block_weights


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

    names = tuple(observable_names)
    blocks_on_gpu = make_run_blocks(block_fn=block_fn, sys=sys, params=params, trial_ops=trial_ops,
                                    meas_ops=meas_ops, prop_ops=prop_ops, observable_names=names)

    def run_blocks(st, n):  # one host transfer per chunk of blocks, as the reference's jitted chunks
        st, scal, obs = blocks_on_gpu(st, ham_data=ham_data, trial_data=trial_data, meas_ctx=meas_ctx,
                                      prop_ctx=prop_ctx, n_blocks=n)
        e = scal["energy"].cpu().numpy()
        w = scal["weight"].cpu().numpy()
        o = {name: v.cpu().numpy() for name, v in zip(names, obs) if v is not None}
        return st, e, w, o

    t0 = time.perf_counter()
    block_e_eq = [float(state.e_estimate.item())]
    block_w_eq = [float(state.weights.sum().item())]
    obs_eq = {name: [] for name in names}
    log("\nEquilibration:\n")
    chunk = max(1, params.n_eql_blocks // 5)
    for start in range(0, params.n_eql_blocks, chunk):
        n = min(chunk, params.n_eql_blocks - start)
        state, e, w, o = run_blocks(state, n)
        block_e_eq.extend(e.tolist())
        block_w_eq.extend(w.tolist())
        for name in names:
            obs_eq[name].append(o[name])
        log(f"[eql {start + n:4d}/{params.n_eql_blocks}]  {float(np.mean(e * w) / np.mean(w)):14.10f}  "
            f"{float(np.mean(w)):12.6e}  {int(state.node_encounters.item()):10d}  {time.perf_counter() - t0:8.1f}")

    log("\nSampling:\n")
    target_error = 0.0 if target_error is None else target_error
    block_e_s: list[float] = []
    block_w_s: list[float] = []
    obs_s = {name: [] for name in names}
    chunk = max(1, params.n_blocks // 10)
    t_mark = time.perf_counter()
    for start in range(0, params.n_blocks, chunk):
        n = min(chunk, params.n_blocks - start)
        state, e, w, o = run_blocks(state, n)
        block_e_s.extend(e.tolist())
        block_w_s.extend(w.tolist())
        for name in names:
            obs_s[name].append(o[name])
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
    if verbose:
        log("\nFinal blocking analysis:")
    stats = blocking_analysis_ratio(clean[:, 0], clean[:, 1], print_q=verbose)
    w_s = clean[:, 1]
    b_star = stats["B_star"]
    block_obs, means, errs = {}, {}, {}
    for name in names:  # trot/driver.py:352-400
        eq = np.concatenate(obs_eq[name], axis=0) if obs_eq[name] else None
        smp = np.concatenate(obs_s[name], axis=0)[keep] if obs_s[name] else None
        if eq is None and smp is None:
            block_obs[name] = np.zeros((0,))
            means[name] = np.zeros((0,))
            errs[name] = np.zeros((0,))
            continue
        block_obs[name] = np.concatenate([eq if eq is not None else smp[:0], smp if smp is not None else eq[:0]])
        if smp is None:
            means[name] = np.zeros((0,))
            errs[name] = np.zeros((0,))
            continue
        means[name] = weighted_block_mean(smp, w_s)
        errs[name] = np.full(smp.shape[1:], np.nan)
        if b_star is not None and b_star >= 1:
            num, den = rebin_observable(smp, w_s, b_star)
            if num.shape[0] >= 2:
                errs[name] = jackknife_ratios(num, den)[1]
    return QmcResult(
        mean_energy=stats["mu"],
        stderr_energy=stats["se_star"],
        block_energies=np.concatenate([block_e_eq, clean[:, 0]]),
        block_weights=np.concatenate([block_w_eq, clean[:, 1]]),
        block_observables=block_obs,
        observable_means=means,
        observable_stderrs=errs,
    )


def weighted_block_mean(values: np.ndarray, weights: np.ndarray) -> np.ndarray:
    """trot/driver.py:56-61: sum_b w_b values_b / sum_b w_b over blocks, 0 when the weights sum to 0."""
    values = np.asarray(values)
    weights = np.asarray(weights).ravel()
    num = np.tensordot(weights, values, axes=(0, 0))
    return np.zeros_like(num) if weights.sum() == 0 else num / weights.sum()
