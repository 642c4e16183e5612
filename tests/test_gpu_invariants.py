"""Data-parallel invariants of the step: the reference's sharded == unsharded tests (tests/test_sharding.py
:49-228) restated for one GPU.

* Walker chunks: the propagation run in chunks of 7 walkers (``make_prop_ops(walker_chunk=7)``) gives the
  same bits as all walkers at once, FP64 and split-BF16 (every walker's VHS row, Taylor series and
  one-body products are computed by the same kernels in the same K order whatever the chunk).
* Walker order: permuting the walkers (with their fields and weights) permutes the walkers, overlaps and
  weights after a step in the same way, bit for bit; the population shift, a sum over walkers, agrees to
  rounding; so does the block energy.
* Walker kind (tests/test_uhf.py:131-170): restricted walkers w and unrestricted walkers (w, w) agree to
  1e-12 in every measurement and after two steps.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.cases import problem, rel

pytestmark = pytest.mark.gpu


def _setup(p, precision, chunk):
    from trot_amd import meas, prop, trial
    from trot_amd.core import HamChol, System

    N, n = p["h1"].shape[0], p["ca"].shape[1]
    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    tops = trial.make_rhf_trial_ops(sys_)
    mops = meas.make_rhf_meas_ops(sys_, precision=precision)
    pops = prop.make_prop_ops("restricted", "restricted", precision=precision, taylor_precision=precision,
                              walker_chunk=chunk, taylor_hi_from=3 if precision == "bf16x3" else 0)
    W = p["wa"].shape[0]
    params = prop.QmcParams(dt=0.01, n_walkers=W, seed=3, n_prop_steps=3)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    kw = dict(params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
              prop_ctx=pctx)
    return sys_, pops, kw


def _init(sys_, pops, kw, w):
    return pops.init_prop_state(sys=sys_, ham_data=kw["ham_data"], trial_ops=kw["trial_ops"],
                                trial_data=kw["trial_data"], meas_ops=kw["meas_ops"], params=kw["params"],
                                initial_walkers=torch.from_numpy(np.ascontiguousarray(w)).cuda())


@pytest.mark.parametrize("precision", ["f64", "bf16x3"])
def test_walker_chunks_give_the_same_bits(gpu, precision):
    print(f"[invariants] chunks {precision}", flush=True)
    p = problem(40, 6, 50, 30, seed=61)
    out = []
    for chunk in (0, 7):
        sys_, pops, kw = _setup(p, precision, chunk)
        st = _init(sys_, pops, kw, p["wa"])
        for s in range(2):
            st = pops.step(st, fields=torch.from_numpy(p["fields"][s]).cuda(), **kw)
        out.append(st)
    a, b = out
    assert torch.equal(a.walkers, b.walkers)
    assert torch.equal(a.weights, b.weights) and torch.equal(a.overlaps, b.overlaps)
    assert torch.equal(a.pop_control_ene_shift, b.pop_control_ene_shift)


def test_walker_order_is_immaterial(gpu):
    print("[invariants] permutation", flush=True)
    p = problem(32, 5, 40, 24, seed=62)
    perm = np.random.default_rng(0).permutation(24)
    sys_, pops, kw = _setup(p, "f64", 0)
    st_a = _init(sys_, pops, kw, p["wa"])
    st_b = _init(sys_, pops, kw, p["wa"][perm])
    # the same initial energy estimate for both (the reference takes it from walker 0)
    st_b = st_b._replace(e_estimate=st_a.e_estimate.clone(),
                         pop_control_ene_shift=st_a.pop_control_ene_shift.clone())
    pt = torch.from_numpy(perm).cuda()
    for s in range(3):
        f = p["fields"][s]
        st_a = pops.step(st_a, fields=torch.from_numpy(f).cuda(), **kw)
        st_b = pops.step(st_b, fields=torch.from_numpy(np.ascontiguousarray(f[perm])).cuda(), **kw)
        assert torch.equal(st_a.walkers[pt], st_b.walkers), s
        assert torch.equal(st_a.overlaps[pt], st_b.overlaps), s
        # weights depend on the population shift, a sum over walkers: equal to rounding
        assert rel(st_b.weights.cpu().numpy(), st_a.weights[pt].cpu().numpy()) < 1e-13, s
        assert abs(st_a.pop_control_ene_shift.item() - st_b.pop_control_ene_shift.item()) < 1e-12 * max(
            1.0, abs(st_a.pop_control_ene_shift.item())), s
    e_a = kw["meas_ops"].engine.energy_from_green(kw["meas_ops"].engine.overlap_green(st_a.walkers,
                                                                                       kw["trial_data"])[1],
                                                   float(p["h0"]), kw["meas_ctx"])
    e_b = kw["meas_ops"].engine.energy_from_green(kw["meas_ops"].engine.overlap_green(st_b.walkers,
                                                                                       kw["trial_data"])[1],
                                                   float(p["h0"]), kw["meas_ctx"])
    assert torch.equal(e_a[pt], e_b)  # per-walker energies: the same kernels on the same walkers
    wa, wb = st_a.weights, st_b.weights
    blk_a = complex((wa * e_a).sum().item() / wa.sum().item())
    blk_b = complex((wb * e_b).sum().item() / wb.sum().item())
    assert abs(blk_a - blk_b) < 1e-12 * abs(blk_a)


def test_restricted_walkers_equal_unrestricted_with_equal_spins(gpu):
    """The reference's walker-kind equivalence (tests/test_uhf.py:131-170, at 1e-12): restricted walkers w and
    unrestricted walkers (w, w) under the RHF trial give the same overlaps, force bias, local energy and
    observables, and the same walkers, weights and population shift after two steps."""
    from trot_amd import meas, prop, trial
    from trot_amd.core import HamChol, System

    print("[invariants] restricted == unrestricted (w, w)", flush=True)
    p = problem(24, 5, 30, 12, seed=63)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    out = {}
    for kind in ("restricted", "unrestricted"):
        sys_ = System(24, (5, 5), kind)
        mops = meas.make_rhf_meas_ops(sys_)
        tops = trial.make_rhf_trial_ops(sys_)
        pops = prop.make_prop_ops("restricted", kind)
        params = prop.QmcParams(dt=0.01, n_walkers=12, seed=3)
        pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
        mctx = mops.build_meas_ctx(ham, tr)
        w = torch.from_numpy(p["wa"]).cuda()
        wk = w if kind == "restricted" else (w, w.clone())
        res = {"ov": tops.overlap(wk, tr), "fb": mops.require_kernel("force_bias")(wk, ham, mctx, tr),
               "e": mops.require_kernel("energy")(wk, ham, mctx, tr),
               "dm": mops.require_observable("rdm1")(wk, ham, mctx, tr),
               "dc": mops.require_observable("density_corr")(wk, ham, mctx, tr)}
        st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                  params=params, initial_walkers=wk)
        for s in range(2):
            st = pops.step(st, fields=torch.from_numpy(p["fields"][s]).cuda(), params=params, ham_data=ham,
                           trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx, prop_ctx=pctx)
        res["state"] = st
        out[kind] = res
    r, u = out["restricted"], out["unrestricted"]
    for name in ("ov", "fb", "e", "dm", "dc"):
        assert rel(u[name].cpu().numpy(), r[name].cpu().numpy()) < 1e-12, name
    sr, su = r["state"], u["state"]
    for x in su.walkers:
        assert rel(x.cpu().numpy(), sr.walkers.cpu().numpy()) < 1e-12
    assert rel(su.weights.cpu().numpy(), sr.weights.cpu().numpy()) < 1e-12
    assert rel(su.overlaps.cpu().numpy(), sr.overlaps.cpu().numpy()) < 1e-12
    assert abs(su.pop_control_ene_shift.item() - sr.pop_control_ene_shift.item()) < 1e-12 * abs(
        sr.pop_control_ene_shift.item())
