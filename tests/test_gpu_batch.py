"""The batching form: 32 fragments of uneven size advanced together through one HIP graph per step.

Fragments as the batching case defines them (trot_amd.synthetic.fragment_sizes: N ~ U[40, 200] with
seed 4, n = round(N/5), G = 4N), each a synthetic system by the scale-up recipe (seed 100 + i, tau 256)
with a UHF trial of (n, n - 1) occupied orbitals and 128 complex128 walkers, dt = 0.005.  The same 100
steps with the same pre-drawn fields run once fragment by fragment (eager launches) and once through
FragmentBatch (one captured graph replayed per step: grouped launches over all fragments, with the whole
propagation of every fragment in one fused launch or as one-body / Taylor-term launches with autotuned
tile shapes, or the fragments on four parallel branches); walkers, weights, overlaps, population shifts
and node counts must be bit-identical after steps 1, 10 and 100.

Mixed precision: the same 32 fragments with split-BF16 GEMMs (Taylor terms 3-6 one bf16 pass), through one
graph of grouped launches (every launch of the split-BF16 propagation over all walker sets) and through the
graph of per-fragment launches on four branches, each bit-identical to the one-by-one eager mixed steps after
20 steps.

The fused propagation alone (trot_fused_prop_grouped) is also checked bit for bit against
trot_one_body_apply, trot_taylor_apply_symblk and trot_one_body_apply on walker sets from N = 9, n = 1 to
N = 256, n = 32 (sixteen 16-row tiles, the largest operand that fits in LDS) in one launch.
"""

from __future__ import annotations

import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

W, DT, STEPS = 128, 0.005, 100


def _fragments(n_frag=32, precision="f64"):
    from trot_amd import batch, meas, prop, synthetic, trial
    from trot_amd.core import HamChol, System

    out = []
    for i, (N, n, G) in enumerate(synthetic.fragment_sizes(n_frag, seed=4)):
        syn = synthetic.make_synthetic(N, n, G, W, 100 + i, 256.0, nocc_b=n - 1)
        sys_ = System(N, (n, n - 1), "unrestricted")
        ham = HamChol(syn.h0, syn.h1, syn.chol).to_device()
        tr = trial.UhfTrial(syn.trial, syn.trial_b)
        mops = meas.make_uhf_meas_ops(sys_, precision=precision)
        tops = trial.make_uhf_trial_ops(sys_)
        if precision == "f64":
            pops = prop.make_prop_ops("restricted", "unrestricted")
        else:  # the headline's mixed form: split-BF16 GEMMs, Taylor terms 3-6 one bf16 pass
            pops = prop.make_prop_ops("restricted", "unrestricted", precision=precision, taylor_precision=precision,
                                      taylor_hi_from=3)
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


@pytest.mark.parametrize("grouped,fused", [(True, True), (True, False), (False, False)])
def test_fragment_batch_bit_identical(gpu, grouped, fused):
    from trot_amd import batch, synthetic

    print(f"[batch] grouped launches: {grouped}, fused propagation: {fused}", flush=True)
    frags = _fragments()
    init = [f.state for f in frags]
    fb = batch.FragmentBatch(frags, grouped=grouped, fused=fused)
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
    if grouped:
        assert fb.plan.n_fused == (2 * len(frags) if fused else 0)
    print(f"[batch] grouped {grouped} fused {fused}: 32 fragments x {STEPS} steps: eager one-by-one {t_eager:.2f} s, "
          f"graph {t_graph:.2f} s for {STEPS - 1} steps ({1e3 * t_graph / (STEPS - 1):.2f} ms/step); "
          f"fragments with finite weights after {STEPS} steps: {n_finite}; Taylor tiles "
          f"{sorted(set(tiles.values()))} over {len(tiles)} shapes", flush=True)
    for key, a in ref.items():
        b = got[key]
        for x, y in zip(a, b):
            assert np.array_equal(x, y, equal_nan=True), key
    assert all(np.all(np.isfinite(ref[(i, 1)][-4])) for i in range(len(frags)))


def test_fused_propagation_bit_identical(gpu):
    from trot_amd import _lib

    shapes = [(3, 9, 1), (5, 40, 8), (4, 123, 25), (2, 256, 32), (3, 200, 40), (6, 64, 48)]
    g = _lib.Grouped("fused")
    runs = []
    sqrt_dt, K = 0.07, 6
    for s, (W, N, n) in enumerate(shapes):
        gen = torch.Generator().manual_seed(s)
        sym = lambda: (lambda a: (a + a.transpose(1, 2)) / N)(torch.randn((W, N, N), generator=gen,
                                                                           dtype=torch.float64))
        v = torch.complex(_lib.pack_symblk(sym().cuda()), _lib.pack_symblk(sym().cuda())).contiguous()
        e = torch.linalg.matrix_exp(-0.01 * sym()[0]).contiguous().cuda()
        w = torch.randn((W, N, n, 2), generator=gen, dtype=torch.float64)
        w = torch.view_as_complex(w).contiguous().cuda()
        ref = _lib.one_body_apply(e, _lib.taylor_apply_symblk(v, _lib.one_body_apply(e, w), sqrt_dt, K))
        wf, acc = w.clone(), torch.empty_like(w)
        assert g.add_fused(v, e, wf, acc, W, N, n, sqrt_dt, K)
        runs.append((W, N, n, ref, wf))
    # operands beyond LDS (N x n padded, complex FP64, over 160 KiB) or beyond three 16-column tiles are refused
    big = torch.empty((1, 256, 48), dtype=torch.complex128, device="cuda")
    vb = torch.empty((1, _lib.symblk_cols(256)), dtype=torch.complex128, device="cuda")
    eb = torch.empty((256, 256), dtype=torch.float64, device="cuda")
    assert not g.add_fused(vb, eb, big, torch.empty_like(big), 1, 256, 48, sqrt_dt, K)
    wide = torch.empty((1, 64, 49), dtype=torch.complex128, device="cuda")
    assert not g.add_fused(vb, eb, wide, torch.empty_like(wide), 1, 64, 49, sqrt_dt, K)
    print(f"[fused] {len(shapes)} walker sets, largest LDS {g.max_lds} bytes", flush=True)
    g.finish().launch()
    torch.cuda.synchronize()
    for W, N, n, ref, wf in runs:
        assert torch.equal(wf, ref), (W, N, n)
        assert torch.isfinite(torch.view_as_real(wf)).all()


@pytest.mark.parametrize("grouped", [True, False])
def test_mixed_precision_fragments_bit_identical(gpu, grouped):
    from trot_amd import batch, synthetic

    steps = 20
    print(f"[batch] split-BF16 fragments, grouped launches: {grouped}", flush=True)
    frags = _fragments(precision="bf16x3")
    init = [f.state for f in frags]
    fb = batch.FragmentBatch(frags, grouped=grouped)
    assert fb.mixed and fb.grouped == grouped
    fields = [[synthetic.fields(100 + i, s, W, f.n_fields) for i, f in enumerate(frags)] for s in range(steps)]
    ref = []
    for i, f in enumerate(frags):
        st = f.state
        for s in range(steps):
            f.state = st
            st = f.step(torch.from_numpy(fields[s][i]).cuda())
        ref.append(_snapshot(st))
    for f, st in zip(frags, init):
        f.state = st
    packed = [fb.pack_fields(fields[s]) for s in range(steps)]
    fb.capture(packed[0])
    for s in range(1, steps):
        fb.step(packed[s])
    for i, f in enumerate(frags):
        for x, y in zip(ref[i], _snapshot(f.state)):
            assert np.array_equal(x, y, equal_nan=True), i
    assert all(np.all(np.isfinite(r[-4])) for r in ref)
