"""Free-projection AFQMC on the GPU: step, block and the trajectory driver.

Reference: trot/prop/afqmc_fp.py:18-120 (step and factory), trot/prop/chol_afqmc_ops_fp.py:49-74
(context), trot/prop/blocks.py:415-490 (block_fp), trot/driver.py:806-936 (run_qmc_fp),
trot/walkers.py:137-226 (orthogonalize, qr_norm, multiply_constants), trot/prop/types.py:55-62.

The step reuses the phaseless path's kernels for the VHS, the one-body factors and the Taylor
series (ten terms, as the reference hard-codes), and adds the free-projection pieces as HIP
kernels: the walker constant exp(-sqrt(dt) x.mf + dt h0_prop) spread over the orbitals
(trot_fp_prepare, trot_scale_rows), the QR norm |det R| from the Cholesky factor of w^H w
(trot_qr_norm), and the comb with the norm divided back out (trot_sr_comb, trot_fp_unnorm).
"""

from __future__ import annotations

import dataclasses
import time
from dataclasses import dataclass
from typing import Callable

import numpy as np
import torch

from . import _lib
from . import walkers as wk
from .blocks import BlockObs
from .core import MeasOps, System
from .meas import _h0
from .prop import PropOps, PropState, QmcParamsBase, apply_trotter, build_prop_ctx, init_prop_state

FP_TAYLOR_TERMS = 10  # trot/prop/afqmc_fp.py:40


@dataclass(frozen=True)
class QmcParamsFp(QmcParamsBase):
    """trot/prop/types.py:55-62."""

    dt: float = 0.05
    n_prop_steps: int = 20
    n_qr_blocks: int = 10
    n_blocks: int = 5
    ene0: float | None = None
    n_traj: int = 10


def build_prop_ctx_fp(ham_data, rdm1, dt: float, ene0: float, **kw):
    """trot/prop/chol_afqmc_ops_fp.py:49-74: the phaseless context with h0_prop = -h0 - 1/2 sum mf^2 + ene0."""
    ctx = build_prop_ctx(ham_data, rdm1, dt, **kw)
    return dataclasses.replace(ctx, h0_prop=complex(ctx.h0_prop) + float(ene0))


def _spins(w):
    return w if isinstance(w, tuple) else (w,)


def afqmc_step_fp(state: PropState, sys: System, *, params, ham_data, trial_data, meas_ops: MeasOps, prop_ctx,
                  meas_ctx, fields: torch.Tensor | None = None, zeta=None) -> PropState:
    """trot/prop/afqmc_fp.py:18-72.  ``fields`` (W, G) and ``zeta`` override the provider's draws."""
    kind = sys.walker_kind.lower()
    if kind not in ("restricted", "unrestricted"):
        raise ValueError(f"multiply_constants does not handle '{sys.walker_kind}' walkers")
    nw = wk.n_walkers(state.walkers)
    G = prop_ctx.chol_flat.shape[0]
    key = state.rng_key
    if fields is None:
        fields, key = prop_ctx.fields(key, nw, G, state.weights.device)  # afqmc_fp.py:30-32
    else:
        key = (key[0], key[1] + 1)
    spins = _spins(state.walkers)
    n_total = sum(s.shape[-1] for s in spins) if kind == "unrestricted" else 2 * spins[0].shape[-1]
    xbar, _, scale = _lib.fp_prepare(fields.contiguous(), prop_ctx.mf_shifts, prop_ctx.sqrt_dt, prop_ctx.dt,
                                     prop_ctx.h0_prop, n_total)  # afqmc_fp.py:35-36
    w_new = apply_trotter(state.walkers, xbar, prop_ctx, FP_TAYLOR_TERMS)  # afqmc_fp.py:38-40
    new_spins = _spins(w_new)
    for s in new_spins:  # multiply_constants, afqmc_fp.py:42
        _lib.scale_rows(s, scale)
    dets = [_lib.qr_norm(s) for s in new_spins]  # qr_norm, afqmc_fp.py:51
    norms, weights_new = _lib.fp_weights(state.weights, dets[0], dets[1] if kind == "unrestricted" else None)
    if zeta is None:
        zeta, key = prop_ctx.fields.uniform(key)  # afqmc_fp.py:53-54
    else:
        key = (key[0], key[1] + 1)
    idx, weights_sr = _lib.sr_comb(weights_new, zeta)  # afqmc_fp.py:59-60
    walkers_sr = wk.gather(w_new, idx)
    _lib.fp_unnorm(weights_sr, norms, idx)  # afqmc_fp.py:61-62
    return PropState(walkers=walkers_sr, weights=weights_sr, overlaps=state.overlaps, rng_key=key,
                     pop_control_ene_shift=state.pop_control_ene_shift, e_estimate=state.e_estimate,
                     node_encounters=state.node_encounters, green=None)


def make_prop_ops_fp(ham_basis: str, walker_kind: str, sys: System, mixed_precision: bool = False, *,
                     precision: str | None = None, taylor_precision: str | None = None,
                     taylor_hi_from: int = 0) -> PropOps:
    """trot/prop/afqmc_fp.py:75-120.  ``mixed_precision`` runs the VHS on FP32 MFMA (the reference casts
    chol_flat to float32 and xbar to complex64) and the Taylor series in FP64 on V cast back to complex128;
    ``precision`` / ``taylor_precision`` name the forms of make_prop_ops."""
    if ham_basis != "restricted":
        raise NotImplementedError(f"Not implemented for ham_basis={ham_basis}")
    if walker_kind.lower() not in ("restricted", "unrestricted"):
        raise NotImplementedError(f"Not implemented for walker_kind={walker_kind}")
    prec = precision or ("f32" if mixed_precision else "f64")

    def step_fp(state, *, params, ham_data, trial_data, trial_ops, meas_ops, meas_ctx, prop_ctx, fields=None,
                zeta=None):
        return afqmc_step_fp(state, sys, params=params, ham_data=ham_data, trial_data=trial_data, meas_ops=meas_ops,
                             meas_ctx=meas_ctx, prop_ctx=prop_ctx, fields=fields, zeta=zeta)

    def _build(ham_data, rdm1, params):
        if not isinstance(params, QmcParamsFp):
            raise TypeError("free projection needs QmcParamsFp")
        if params.ene0 is None:
            raise ValueError("ene0 must be set for FP propagation")
        return build_prop_ctx_fp(ham_data, rdm1, params.dt, params.ene0, precision=prec,
                                 taylor_precision=taylor_precision, taylor_hi_from=taylor_hi_from)

    return PropOps(init_prop_state=init_prop_state, build_prop_ctx=_build, step=step_fp)


def draws_per_block(params) -> int:
    """Device-stream draws of one free-projection block: fields and a comb offset per step
    (for trot_amd.graph.GraphBlock(block_fp, draws_per_block))."""
    return 2 * params.n_prop_steps


def block_fp(state: PropState, *, sys: System, params, ham_data, trial_data, trial_ops, meas_ops: MeasOps, meas_ctx,
             prop_ops, prop_ctx, sr_fn: Callable | None = None, observable_names: tuple = (), fields_seq=None,
             zetas=None) -> tuple[PropState, BlockObs]:
    """trot/prop/blocks.py:415-490: n_qr_blocks groups of steps, each followed by a QR that moves the norm
    into the weights; then overlaps, E_L clamped around ene0, and the overlap-weighted block sums.
    ``fields_seq`` / ``zetas`` (n_prop_steps each) replace the provider's draws."""
    for name in observable_names:
        raise KeyError(f"missing requested observable '{name}'. available: []")
    if params.n_prop_steps % params.n_qr_blocks:
        raise ValueError("n_prop_steps must be a multiple of n_qr_blocks")
    kind = sys.walker_kind.lower()
    per = params.n_prop_steps // params.n_qr_blocks
    s = 0
    for _ in range(params.n_qr_blocks):
        for _ in range(per):
            state = prop_ops.step(state, params=params, ham_data=ham_data, trial_data=trial_data,
                                  trial_ops=trial_ops, meas_ops=meas_ops, meas_ctx=meas_ctx, prop_ctx=prop_ctx,
                                  fields=None if fields_seq is None else fields_seq[s],
                                  zeta=None if zetas is None else zetas[s])
            s += 1
        qd = [_lib.orthogonalize(x) for x in _spins(state.walkers)]  # blocks.py:451-458
        q = tuple(o[0] for o in qd) if kind == "unrestricted" else qd[0][0]
        _, weights = _lib.fp_weights(state.weights, qd[0][1], qd[1][1] if kind == "unrestricted" else None)
        state = state._replace(walkers=q, weights=weights)
    engine = meas_ops.engine
    overlaps, greens = engine.overlap_green(state.walkers, trial_data)  # blocks.py:463-466
    e_samples = engine.energy_from_green(greens, _h0(ham_data), meas_ctx)  # blocks.py:467-470
    out = _lib.block_energy_fp(e_samples, state.weights, overlaps, params.dt, params.ene0)  # blocks.py:472-488
    state = state._replace(overlaps=overlaps, green=None)
    obs = BlockObs(scalars={"energy": out[0], "weight": out[1], "overlap": out[2], "abs_overlap": out[3]},
                   observables={})
    return state, obs


def run_qmc_fp(*, sys, params: QmcParamsFp, ham_data, trial_data, meas_ops, trial_ops, prop_ops,
               block_fn: Callable = block_fp, state=None, meas_ctx=None, prop_ctx=None, target_error=None, mesh=None,
               verbose: bool = True):
    """trot/driver.py:806-936: n_traj independent trajectories (seed + i) of n_blocks blocks; running mean,
    error over trajectories and the average sign per imaginary time."""
    from .driver import QmcResult

    if mesh is not None:
        raise NotImplementedError("trot_amd runs on one GPU")
    log = print if verbose else (lambda *a, **k: None)
    if prop_ctx is None:
        prop_ctx = prop_ops.build_prop_ctx(ham_data, trial_ops.get_rdm1(trial_data), params)
    if meas_ctx is None:
        meas_ctx = meas_ops.build_meas_ctx(ham_data, trial_data)
    if state is None:
        state = prop_ops.init_prop_state(sys=sys, ham_data=ham_data, trial_ops=trial_ops, trial_data=trial_data,
                                         meas_ops=meas_ops, params=params)
    t0 = time.perf_counter()
    nb = params.n_blocks
    block_e_all = np.zeros((params.n_traj, nb + 1), dtype=np.complex128)
    block_w_all = np.zeros((params.n_traj, nb + 1), dtype=np.complex128)
    total_sign = np.ones((params.n_traj, nb + 1), dtype=np.complex128)
    mean = err = None
    log("\nSampling:\n")
    for i in range(params.n_traj):
        log("Trajectory count", i + 1)
        log(f"{'tau':^12s}    {'E_avg':^14s}  {'E_err':^13s}  {'sign':>6s}")
        if i > 0:
            params = dataclasses.replace(params, seed=params.seed + i)
            state = prop_ops.init_prop_state(sys=sys, ham_data=ham_data, trial_ops=trial_ops, trial_data=trial_data,
                                             meas_ops=meas_ops, params=params)
        ov = state.overlaps
        block_e_all[i, 0] = complex(state.e_estimate.item())
        block_w_all[i, 0] = float(state.weights.sum().item())
        total_sign[i, 0] = complex((ov.sum() / ov.abs().sum()).item())
        scal = []
        for _ in range(nb):
            state, obs = block_fn(state, sys=sys, params=params, ham_data=ham_data, trial_data=trial_data,
                                  trial_ops=trial_ops, meas_ops=meas_ops, meas_ctx=meas_ctx, prop_ops=prop_ops,
                                  prop_ctx=prop_ctx)
            scal.append(torch.stack([obs.scalars[k] for k in ("energy", "weight", "overlap", "abs_overlap")]))
        sc = torch.stack(scal).cpu().numpy() if scal else np.zeros((0, 4), dtype=np.complex128)
        block_e_all[i, 1:] = sc[:, 0]
        block_w_all[i, 1:] = sc[:, 1]
        total_sign[i, 1:] = sc[:, 2] / sc[:, 3]
        wsum = np.sum(block_w_all[: i + 1], axis=0)
        mean = np.sum(block_e_all[: i + 1] * block_w_all[: i + 1], axis=0) / wsum
        sign = np.sum(total_sign[: i + 1] * block_w_all[: i + 1], axis=0) / wsum
        err = np.zeros_like(mean) if i == 0 else np.std(block_e_all[: i + 1], axis=0) / np.sqrt(i)
        tau = params.dt * params.n_prop_steps * np.arange(nb + 1)
        for j in range(nb + 1):
            log(f"{tau[j]:12.4f}    {mean[j].real:14.10f}  {err[j].real:13.7e}  {sign[j].real:6.2f}")
        log(f"Wall time :{time.perf_counter() - t0:12.1f} s\n")
    return QmcResult(mean_energy=mean, stderr_energy=err, block_energies=block_e_all, block_weights=block_w_all,
                     block_observables={}, observable_means={}, observable_stderrs={})
