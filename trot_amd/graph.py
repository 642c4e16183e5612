"""The block driver on the device: one block (trot/prop/blocks.py:195-287) recorded as one HIP graph.

``GraphBlock`` is a drop-in ``block_fn`` for ``trot_amd.driver.run_qmc`` (the reference's driver takes
its block function the same way, trot/driver.py:64-110).  The first call runs one block eagerly and
records the next one into a HIP graph (torch.cuda.CUDAGraph on the ROCm stream) over static state
buffers; every later call is one graph launch.  The n_prop_steps steps, QR, overlap recompute, local
energy, clamp / EMA, the comb offset draw and SR all stay on the GPU: a block costs one launch and
no host round trip, where the eager block issues about thirty launches per step from Python.

The auxiliary fields and the comb offset come from the counter-based Philox stream of
``trot_amd.prop.DeviceFields``, with the counter held on the device (``StateFields``) so that each
replay draws the next numbers; a replayed block is therefore bit-identical to the eager block of the
same key.  Captured code must not synchronise with the host, allocate outside torch's graph pool or
record memset nodes; the step and block paths of trot_amd satisfy all three.
"""

from __future__ import annotations

import dataclasses

import torch

from . import _lib
from .batch import _static_copy
from .blocks import BlockObs, block as eager_block
from .prop import DeviceFields, PropState, StateFields


_FIELDS = ("walkers", "weights", "overlaps", "pop_control_ene_shift", "e_estimate", "node_encounters", "green")


def _copy_state(dst: PropState, src: PropState) -> None:
    """dst's tensors <- src's, field by field; fields that dst does not hold are skipped."""
    for name in _FIELDS:
        d, s = getattr(dst, name), getattr(src, name)
        if d is None:
            continue
        if s is None:
            raise ValueError(f"the state has no {name}, which the recording holds")
        d = d if isinstance(d, tuple) else (d,)
        s = s if isinstance(s, tuple) else (s,)
        if len(d) != len(s):
            raise ValueError(f"the state's {name} does not match the recorded structure")
        for a, b in zip(d, s):
            if a is None:
                continue
            if b is None or a.shape != b.shape:
                raise ValueError(f"the state's {name} does not match the recorded shapes")
            a.copy_(b)


class _Recorded:
    """Static state buffers, a device counter and one recorded HIP graph of ``run(static) -> (state, scalars, arrays)``.

    The first call runs ``run`` eagerly (on a side stream, as graph capture requires) on a copy of the
    given state; its result fixes the structure of the static buffers and is the first call's result.
    Then ``run`` is recorded on the static buffers, writing back into them."""

    def __init__(self):
        self.graph = None
        self.static = None
        self.last = None
        self.rng = None
        self.obs = None
        self.arrays = ()
        self.replays = 0

    def _store(self, state_out, scalars, arrays=()):
        _copy_state(self.static, state_out)
        if scalars:
            self.obs.copy_(torch.stack([x.reshape(()) for x in scalars]))
        for dst, src in zip(self.arrays, arrays):
            dst.copy_(src)

    def _load(self, state):
        _copy_state(self.static, state)
        self.rng.copy_(_lib.rng_state(*state.rng_key))

    def _device_ctx(self, kw):
        pctx = kw["prop_ctx"]
        if not isinstance(pctx.fields, (DeviceFields, StateFields)):
            raise ValueError("recording needs the device field stream (DeviceFields) on the propagation context")
        out = dict(kw)
        # inside the recording the propagation may overwrite the static walkers it was given (nothing reads
        # them after it), which saves copying the new walkers back into them at the end of every replay
        out["prop_ctx"] = dataclasses.replace(pctx, fields=StateFields(self.rng),
                                              low={**pctx.low, "walkers_in_place": True})
        return out

    def _run(self, state, kw, run, n_draws):
        seed, ctr = state.rng_key
        if self.graph is None:
            self.rng = _lib.rng_state(seed, ctr)
            args = self._device_ctx(kw)
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                out, scalars, arrays = run(_static_copy(state), args)
                self.static = _static_copy(out)
                if scalars:
                    self.obs = torch.stack([x.reshape(()) for x in scalars]).to(torch.complex128 if any(
                        x.is_complex() for x in scalars) else torch.float64).contiguous()
                self.arrays = tuple(a.detach().clone().contiguous() for a in arrays)
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            del out, scalars, arrays
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._store(*run(self.static, args))
            # recording does not run the work: the buffers hold the warm-up's result and the device
            # counter has advanced by the warm-up's draws only
        else:
            if self.last is None or state.walkers is not self.last.walkers:
                self._load(state)
            self.graph.replay()
        self.replays += 1
        self.last = self.static._replace(rng_key=(seed, ctr + n_draws))
        return self.last


class GraphStep(_Recorded):
    """``prop_ops.step`` recorded as one HIP graph: callable with the step's signature (no ``fields``:
    they are drawn on the device).  The first call runs one step eagerly and records the next."""

    def __init__(self, prop_ops, draws_per_step: int = 1):
        """``draws_per_step``: device-stream draws of one step (1 for the phaseless step, 2 for free
        projection: fields and the comb offset)."""
        super().__init__()
        self.prop_ops = prop_ops
        self.draws_per_step = draws_per_step

    def __call__(self, state, **kw):
        if kw.pop("fields", None) is not None:
            raise ValueError("GraphStep draws its fields on the device; fields is not supported")

        def run(st, args):
            return self.prop_ops.step(st, **args), None, ()

        return self._run(state, kw, run, self.draws_per_step)


class GraphBlock(_Recorded):
    """Callable with the signature of ``trot_amd.blocks.block``; records itself on first use.

    The recording is tied to the shapes and contexts of the first call (walker count, Hamiltonian,
    trial, precision).  A state other than the one this object returned last is copied into the
    static buffers before the replay (same shapes required); its ``rng_key`` becomes the device
    counter.  Block scalars are returned as fresh GPU tensors (copied out of the graph's buffers).
    """

    def __init__(self, block_fn=eager_block, draws_per_block=None):
        """``draws_per_block(params)``: device-stream draws one block makes (default: a field draw per step
        and one comb offset, the phaseless block; the free-projection block draws two per step)."""
        super().__init__()
        self.block_fn = block_fn
        self.draws_per_block = draws_per_block or (lambda params: params.n_prop_steps + 1)
        self.names = None
        self.obs_names = ()

    @property
    def blocks_done(self) -> int:
        return self.replays

    def __call__(self, state, **kw):
        if kw.pop("fields_seq", None) is not None:
            raise ValueError("GraphBlock draws its fields on the device; fields_seq is not supported")

        names = tuple(kw.get("observable_names", ()))
        if self.names is not None and names != self.obs_names:
            raise ValueError(f"recorded with observables {self.obs_names}, called with {names}")

        def run(st, args):
            out, obs = self.block_fn(st, **args)
            if self.names is None:
                self.names = list(obs.scalars)
                self.obs_names = tuple(obs.observables)
            return out, tuple(obs.scalars[k] for k in self.names), tuple(obs.observables[k] for k in self.obs_names)

        st = self._run(state, kw, run, self.draws_per_block(kw["params"]))
        obs = BlockObs(scalars={k: self.obs[i].clone() for i, k in enumerate(self.names)},
                       observables={k: self.arrays[i].clone() for i, k in enumerate(self.obs_names)})
        return st, obs
