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
from .batch import _static_copy, _tensors
from .blocks import BlockObs, block as eager_block
from .prop import DeviceFields, StateFields


class _Recorded:
    """Static state buffers, a device counter and one recorded HIP graph of ``run(static) -> (state, scalars)``."""

    def __init__(self):
        self.graph = None
        self.static = None
        self.last = None
        self.rng = None
        self.obs = None
        self.replays = 0

    def _store(self, state_out, scalars):
        for dst, src in zip(_tensors(self.static), _tensors(state_out)):
            dst.copy_(src)
        if scalars is not None:
            self.obs.copy_(torch.stack([x.reshape(()) for x in scalars]))

    def _load(self, state):
        if state.green is None:
            raise ValueError("the state needs its Green's-function cache (PropState.green)")
        dst, new = _tensors(self.static), _tensors(state)
        if len(dst) != len(new) or any(a.shape != b.shape for a, b in zip(dst, new)):
            raise ValueError("the state does not match the recorded shapes")
        for a, b in zip(dst, new):
            a.copy_(b)
        self.rng.copy_(_lib.rng_state(*state.rng_key))

    def _device_ctx(self, kw):
        pctx = kw["prop_ctx"]
        if not isinstance(pctx.fields, (DeviceFields, StateFields)):
            raise ValueError("recording needs the device field stream (DeviceFields) on the propagation context")
        out = dict(kw)
        out["prop_ctx"] = dataclasses.replace(pctx, fields=StateFields(self.rng))
        return out

    def _run(self, state, kw, run, n_draws, n_scalars):
        seed, ctr = state.rng_key
        if self.graph is None:
            if state.green is None:
                _, g = kw["meas_ops"].engine.overlap_green(state.walkers, kw["trial_data"])
                state = state._replace(green=g)
            self.static = _static_copy(state)
            self.rng = _lib.rng_state(seed, ctr)
            self.obs = torch.zeros((max(n_scalars, 1),), dtype=torch.float64, device="cuda")
            args = self._device_ctx(kw)
            # warm-up: eager, on a side stream (graph capture requires it); it is the first call's result
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._store(*run(self.static, args))
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
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

    def __init__(self, prop_ops):
        super().__init__()
        self.prop_ops = prop_ops

    def __call__(self, state, **kw):
        if kw.pop("fields", None) is not None:
            raise ValueError("GraphStep draws its fields on the device; fields is not supported")

        def run(st, args):
            return self.prop_ops.step(st, **args), None

        return self._run(state, kw, run, 1, 0)


class GraphBlock(_Recorded):
    """Callable with the signature of ``trot_amd.blocks.block``; records itself on first use.

    The recording is tied to the shapes and contexts of the first call (walker count, Hamiltonian,
    trial, precision).  A state other than the one this object returned last is copied into the
    static buffers before the replay (same shapes required); its ``rng_key`` becomes the device
    counter.  Block scalars are returned as fresh GPU tensors (copied out of the graph's buffers).
    """

    def __init__(self, block_fn=eager_block):
        super().__init__()
        self.block_fn = block_fn

    @property
    def blocks_done(self) -> int:
        return self.replays

    def __call__(self, state, **kw):
        if kw.pop("fields_seq", None) is not None:
            raise ValueError("GraphBlock draws its fields on the device; fields_seq is not supported")

        def run(st, args):
            out, obs = self.block_fn(st, **args)
            return out, (obs.scalars["energy"], obs.scalars["weight"])

        n_draws = kw["params"].n_prop_steps + 1  # one field draw per step, one comb offset
        st = self._run(state, kw, run, n_draws, 2)
        obs = BlockObs(scalars={"energy": self.obs[0].clone(), "weight": self.obs[1].clone()}, observables={})
        return st, obs
