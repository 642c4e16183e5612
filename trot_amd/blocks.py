"""One propagation + measurement block (trot/prop/blocks.py:195-287) on the GPU, and the per-block state
files of trot/prop/blocks.py:66-192 (``make_block_state_logger`` and its loaders)."""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable

import numpy as np
import torch

from . import _lib
from . import walkers as wk
from .core import MeasOps, System, TrialOps
from .meas import _h0
from .prop import PropState


@dataclass
class BlockObs:
    """trot/prop/blocks.py (BlockObs): scalars {"energy", "weight"} as 0-dim GPU tensors; observables
    {name: weighted block mean} as GPU tensors."""

    scalars: dict
    observables: dict = field(default_factory=dict)


def block(state: PropState, *, sys: System, params, ham_data, trial_data, trial_ops: TrialOps, meas_ops: MeasOps,
          meas_ctx, prop_ops, prop_ctx, sr_fn: Callable = wk.stochastic_reconfiguration,
          observable_names: tuple = (), fields_seq=None) -> tuple[PropState, BlockObs]:
    """Steps, QR, overlap recompute, energy, clamp / weighted mean / EMA, block observables, SR, overlap
    recompute.  Each name of ``observable_names`` is a MeasOps observable; its block value is the mean of the
    per-walker samples weighted by the block's weights (0 when they sum to 0).

    ``fields_seq`` (optional) is a sequence of n_prop_steps (W, G) field tensors
    that replaces the provider, for step-by-step comparisons.
    """
    obs_kernels = {name: meas_ops.require_observable(name) for name in observable_names}
    for s in range(params.n_prop_steps):  # blocks.py:213-228
        f = None if fields_seq is None else fields_seq[s]
        state = prop_ops.step(state, params=params, ham_data=ham_data, trial_data=trial_data, trial_ops=trial_ops,
                              meas_ops=meas_ops, meas_ctx=meas_ctx, prop_ctx=prop_ctx, fields=f)
    engine = meas_ops.engine
    walkers_new = wk.orthonormalize(state.walkers, sys.walker_kind)  # blocks.py:230
    overlaps_new, greens = engine.overlap_green(walkers_new, trial_data)  # blocks.py:231-234
    e_samples = engine.energy_from_green(greens, _h0(ham_data), meas_ctx)  # blocks.py:236-240
    weights = state.weights.clone()
    scal = torch.stack([state.pop_control_ene_shift, state.e_estimate]).to(torch.float64).contiguous()
    out = _lib.block_energy(e_samples, weights, scal, params.dt, params.shift_ema)  # blocks.py:240-257
    obs_means = {}
    for name, kernel in obs_kernels.items():  # blocks.py:259-268, weights with non-finite energies zeroed
        if name in getattr(engine, "observables", ()) and hasattr(engine, "observable_mean_from_green"):
            mean = engine.observable_mean_from_green(name, greens, weights, out[1:2], meas_ctx, trial_data)
            if mean is not None:  # a linear observable: reduced before the product, no per-walker samples
                obs_means[name] = mean
                continue
        if name in getattr(engine, "observables", ()):  # from the Green's functions already at hand
            samples = engine.observable_from_green(name, greens, meas_ctx, trial_data)
        else:
            samples = kernel(walkers_new, ham_data, meas_ctx, trial_data)
        samples = torch.as_tensor(samples, device=weights.device)
        if samples.dtype not in (torch.float64, torch.complex128):
            samples = samples.to(torch.complex128 if samples.is_complex() else torch.float64)
        obs_means[name] = _lib.weighted_mean(samples.contiguous(), weights, out[1:2])
    zeta, key = prop_ctx.fields.uniform(state.rng_key)  # blocks.py:270-271
    if sr_fn is wk.stochastic_reconfiguration:
        idx, weights_sr = wk.sr_indices(weights, zeta)
        w_sr = wk.gather(walkers_new, idx)
        overlaps_sr = wk.gather(torch.view_as_real(overlaps_new).contiguous(), idx)
        overlaps_sr = torch.view_as_complex(overlaps_sr)
        greens_sr = wk.gather(greens, idx)
    else:
        w_sr, weights_sr = sr_fn(walkers_new, weights, zeta, sys.walker_kind)
        w_sr = wk.as_device_walkers(w_sr)
        overlaps_sr, greens_sr = engine.overlap_green(w_sr, trial_data)  # blocks.py:273-276
    state = PropState(
        walkers=w_sr,
        weights=weights_sr,
        overlaps=overlaps_sr,
        rng_key=key,
        pop_control_ene_shift=scal[0],
        e_estimate=scal[1],
        node_encounters=state.node_encounters,
        green=greens_sr,
    )
    return state, BlockObs(scalars={"energy": out[0], "weight": out[1]}, observables=obs_means)


# ---------------------------------------------------------------------------
# per-block state files (trot/prop/blocks.py:66-192)
# ---------------------------------------------------------------------------


def dump_prop_state_npz(state: PropState, path, *, compressed: bool = True) -> None:
    """Write a PropState to one ``.npz`` file (trot/prop/blocks.py:66-98): walkers (``walkers`` or
    ``walkers_0`` / ``walkers_1`` per spin), weights, overlaps, rng_key (seed, counter), the population
    shift, e_estimate and the node count.  The Green's-function cache is not written: it is recomputed
    from the walkers when needed."""
    out = Path(path).expanduser().resolve()
    out.parent.mkdir(parents=True, exist_ok=True)

    def host(x):
        return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)

    arrays = {
        "weights": host(state.weights),
        "overlaps": host(state.overlaps),
        "rng_key": np.asarray([int(k) for k in state.rng_key], dtype=np.int64),
        "pop_control_ene_shift": host(state.pop_control_ene_shift),
        "e_estimate": host(state.e_estimate),
        "node_encounters": host(state.node_encounters),
    }
    if isinstance(state.walkers, tuple):
        for i, w in enumerate(state.walkers):
            arrays[f"walkers_{i}"] = host(w)
    else:
        arrays["walkers"] = host(state.walkers)
    (np.savez_compressed if compressed else np.savez)(out, **arrays)


def load_prop_state_npz(path, device="cuda") -> PropState:
    """A PropState written by ``dump_prop_state_npz`` (trot/prop/blocks.py:101-121), its tensors on
    ``device``; ``green`` is None."""
    with np.load(Path(path).expanduser().resolve()) as data:
        def dev(name):
            return torch.from_numpy(np.array(data[name], order="C")).to(device)  # keeps 0-d arrays 0-d

        if "walkers" in data.files:
            walkers = dev("walkers")
        else:
            walkers = tuple(dev(k) for k in sorted(k for k in data.files if k.startswith("walkers_")))
        return PropState(walkers=walkers, weights=dev("weights"), overlaps=dev("overlaps"),
                         rng_key=tuple(int(k) for k in data["rng_key"]),
                         pop_control_ene_shift=dev("pop_control_ene_shift"), e_estimate=dev("e_estimate"),
                         node_encounters=dev("node_encounters"))


def load_all_prop_states_npz(directory, *, prefix: str = "block_state", device="cuda") -> list:
    """Every ``{prefix}_*.npz`` of ``directory``, in file-name order (trot/prop/blocks.py:124-131)."""
    files = sorted(Path(directory).expanduser().resolve().glob(f"{prefix}_*.npz"))
    return [load_prop_state_npz(f, device) for f in files]


def make_block_state_logger(directory, *, base_block_fn: Callable | None = None, prefix: str = "block_state",
                            start_index: int = 0, compressed: bool = True) -> Callable:
    """A block function that runs ``base_block_fn`` (default ``block``) and writes the state at the end
    of every block to ``{directory}/{prefix}_{index:05d}.npz`` (trot/prop/blocks.py:134-192).  Each file
    is written after the block's kernels finish (a device-to-host copy per block)."""
    base = block if base_block_fn is None else base_block_fn
    out_dir = Path(directory).expanduser().resolve()
    out_dir.mkdir(parents=True, exist_ok=True)
    counter = [int(start_index)]

    def logging_block_fn(state: PropState, **kw):
        state, obs = base(state, **kw)
        dump_prop_state_npz(state, out_dir / f"{prefix}_{counter[0]:05d}.npz", compressed=compressed)
        counter[0] += 1
        return state, obs

    return logging_block_fn
