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


class GraphBlock:
    """Callable with the signature of ``trot_amd.blocks.block``; records itself on first use.

    The recording is tied to the shapes and contexts of the first call (walker count, Hamiltonian,
    trial, precision).  A state other than the one this object returned last is copied into the
    static buffers before the replay (same shapes required); its ``rng_key`` becomes the device
    counter.  Block scalars are returned as fresh GPU tensors (copied out of the graph's buffers).
    """

    def __init__(self, block_fn=eager_block):
        self.block_fn = block_fn
        self.graph = None
        self.static = None
        self.last = None
        self.rng = None
        self.obs = None
        self.blocks_done = 0

    # -- helpers -----------------------------------------------------------
    def _kw(self, kw, pctx):
        out = dict(kw)
        out["prop_ctx"] = pctx
        return out

    def _store(self, state_out, obs):
        for dst, src in zip(_tensors(self.static), _tensors(state_out)):
            dst.copy_(src)
        self.obs.copy_(torch.stack([obs.scalars["energy"].reshape(()), obs.scalars["weight"].reshape(())]))

    def _load(self, state):
        src = state
        if src.green is None:
            raise ValueError("GraphBlock: the state needs its Green's-function cache (PropState.green)")
        dst = _tensors(self.static)
        new = _tensors(src)
        if len(dst) != len(new) or any(a.shape != b.shape for a, b in zip(dst, new)):
            raise ValueError("GraphBlock: the state does not match the recorded shapes")
        for a, b in zip(dst, new):
            a.copy_(b)
        self.rng.copy_(_lib.rng_state(*state.rng_key))

    def _result(self, key):
        self.last = self.static._replace(rng_key=key)
        obs = BlockObs(scalars={"energy": self.obs[0].clone(), "weight": self.obs[1].clone()}, observables={})
        return self.last, obs

    # -- the block -----------------------------------------------------------
    def __call__(self, state, **kw):
        if kw.get("fields_seq") is not None:
            raise ValueError("GraphBlock draws its fields on the device; fields_seq is not supported")
        kw.pop("fields_seq", None)
        pctx = kw["prop_ctx"]
        params = kw["params"]
        if not isinstance(pctx.fields, (DeviceFields, StateFields)):
            raise ValueError("GraphBlock needs the device field stream (DeviceFields) on the propagation context")
        n_draws = params.n_prop_steps + 1  # one field draw per step, one comb offset
        if self.graph is None:
            return self._capture(state, kw, n_draws)
        if self.last is None or state.walkers is not self.last.walkers:
            self._load(state)
        self.graph.replay()
        self.blocks_done += 1
        seed, ctr = state.rng_key
        return self._result((seed, ctr + n_draws))

    def _capture(self, state, kw, n_draws):
        if state.green is None:
            engine = kw["meas_ops"].engine
            _, g = engine.overlap_green(state.walkers, kw["trial_data"])
            state = state._replace(green=g)
        seed, ctr = state.rng_key
        self.static = _static_copy(state)
        self.rng = _lib.rng_state(seed, ctr)
        self.obs = torch.zeros((2,), dtype=torch.float64, device="cuda")
        pctx = dataclasses.replace(kw["prop_ctx"], fields=StateFields(self.rng))
        args = self._kw(kw, pctx)

        # warm-up: the first block, eager, on a side stream (graph capture requires it)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            out, obs = self.block_fn(self.static, **args)
            self._store(out, obs)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        del out, obs

        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            out, obs = self.block_fn(self.static, **args)
            self._store(out, obs)
        # recording does not run the block: the buffers hold the warm-up block's result and the
        # device counter has advanced by that block's draws only
        self.blocks_done = 1
        return self._result((seed, ctr + n_draws))
