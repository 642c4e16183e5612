"""The batching form: 32 fragments of uneven size advanced together through one HIP graph per step.

Fragments as the batching case defines them (trot_amd.synthetic.fragment_sizes: N ~ U[40, 200] with
seed 4, n = round(N/5), G = 4N), each a synthetic system by the scale-up recipe (seed 100 + i, tau 256)
with a UHF trial of (n, n - 1) occupied orbitals and 128 complex128 walkers, dt = 0.005.  The same 100
steps with the same pre-drawn fields run once fragment by fragment (eager launches) and once through
FragmentBatch (one captured graph replayed per step, fragments on four parallel branches, Taylor tile
shapes autotuned per fragment shape); walkers, weights, overlaps, population shifts and node counts must
be bit-identical after steps 1, 10 and 100.
"""

from __future__ import annotations

import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

W, DT, STEPS = 128, 0.005, 100


def _fragments(n_frag=32):
    from trot_amd import batch, meas, prop, synthetic, trial
    from trot_amd.core import HamChol, System

    out = []
    for i, (N, n, G) in enumerate(synthetic.fragment_sizes(n_frag, seed=4)):
        syn = synthetic.make_synthetic(N, n, G, W, 100 + i, 256.0, nocc_b=n - 1)
        sys_ = System(N, (n, n - 1), "unrestricted")
        ham = HamChol(syn.h0, syn.h1, syn.chol).to_device()
        tr = trial.UhfTrial(syn.trial, syn.trial_b)
        mops = meas.make_uhf_meas_ops(sys_)
        tops = trial.make_uhf_trial_ops(sys_)
        pops = prop.make_prop_ops("restricted", "unrestricted")
        params = prop.QmcParams(dt=DT, n_walkers=W, seed=100 + i)
        pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
        mctx = mops.build_meas_ctx(ham, tr)
        w0 = (torch.from_numpy(syn.walkers).cuda(), torch.from_numpy(syn.walkers_b).cuda())
        st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                  params=params, initial_walkers=w0)
        out.append(batch.Fragment(sys=sys_, ham=ham, trial=tr, trial_ops=tops, meas_ops=mops, prop_ops=pops,
                                  params=params, prop_ctx=pctx, meas_ctx=mctx, state=st))
    return out


def _snapshot(st):
    w = st.walkers
    return [x.cpu().numpy() for x in (w if isinstance(w, tuple) else (w,))] + [
        st.weights.cpu().numpy(), st.overlaps.cpu().numpy(), st.pop_control_ene_shift.cpu().numpy(),
        st.node_encounters.cpu().numpy()]


@pytest.mark.parametrize("grouped", [True, False])
def test_fragment_batch_bit_identical(gpu, grouped):
    from trot_amd import batch, synthetic

    print(f"[batch] grouped launches: {grouped}", flush=True)
    frags = _fragments()
    init = [f.state for f in frags]
    fb = batch.FragmentBatch(frags, grouped=grouped)
    tiles = fb.autotune()  # per (walkers, N, n): one of the four FP64 Taylor tile shapes, timed on the GPU
    fields = [[synthetic.fields(100 + i, s, W, f.n_fields) for i, f in enumerate(frags)] for s in range(STEPS)]
    checkpoints = {1, 10, STEPS}

    # one fragment at a time, eager
    ref = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, f in enumerate(frags):
        st = f.state
        for s in range(STEPS):
            f.state = st
            st = f.step(torch.from_numpy(fields[s][i]).cuda())
            if s + 1 in checkpoints:
                ref[(i, s + 1)] = _snapshot(st)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0

    # all fragments through one graph per step
    for f, st in zip(frags, init):
        f.state = st
    packed = [fb.pack_fields(fields[s]) for s in range(STEPS)]
    fb.capture(packed[0])
    got = {(i, 1): _snapshot(f.state) for i, f in enumerate(frags)}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(1, STEPS):
        fb.step(packed[s])
        if s + 1 in checkpoints:
            for i, f in enumerate(frags):
                got[(i, s + 1)] = _snapshot(f.state)
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    n_finite = sum(bool(np.all(np.isfinite(ref[(i, STEPS)][-4]))) for i in range(len(frags)))
    print(f"[batch] grouped {grouped}: 32 fragments x {STEPS} steps: eager one-by-one {t_eager:.2f} s, "
          f"graph {t_graph:.2f} s for {STEPS - 1} steps ({1e3 * t_graph / (STEPS - 1):.2f} ms/step); "
          f"fragments with finite weights after {STEPS} steps: {n_finite}; Taylor tiles "
          f"{sorted(set(tiles.values()))} over {len(tiles)} shapes", flush=True)
    for key, a in ref.items():
        b = got[key]
        for x, y in zip(a, b):
            assert np.array_equal(x, y, equal_nan=True), key
    assert all(np.all(np.isfinite(ref[(i, 1)][-4])) for i in range(len(frags)))
