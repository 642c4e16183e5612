"""Step-by-step and block parity of the FP64 GPU path against the CPU oracle.

Identical Hamiltonian, trial, walkers and pre-drawn fields on both sides;
walkers, weights, overlaps, population shift and node count must agree at
1e-10 relative after every step (the FP64 bar of the phaseless step).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import afqmc_oracle as O
from tests.cases import problem, rel, rel_each

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _setup(p, kind: str, dt: float, uhf: bool = False, layout: str = "auto"):
    from trot_amd import meas, prop, trial
    from trot_amd.core import HamChol, System

    N = p["h1"].shape[0]
    na = p["ca"].shape[1]
    nb = p["cb"].shape[1] if "cb" in p else na
    sys = System(N, (na, nb), kind)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    if uhf:
        tr = trial.UhfTrial(p["ca"], p["cb"])
        mops = meas.make_uhf_meas_ops(sys)
        tops = trial.make_uhf_trial_ops(sys)
        ref_ops = O.make_uhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], p["cb"])
        rdm1 = O.uhf_rdm1(p["ca"], p["cb"])
    else:
        tr = trial.RhfTrial(p["ca"])
        mops = meas.make_rhf_meas_ops(sys)
        tops = trial.make_rhf_trial_ops(sys)
        ref_ops = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], kind)
        rdm1 = O.rhf_rdm1(p["ca"])
    pops = prop.make_prop_ops("restricted", kind, vhs_layout=layout)
    params = prop.QmcParams(dt=dt, n_walkers=p["wa"].shape[0], seed=3, n_prop_steps=4)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    ref_ctx = O.build_prop_ctx(p["h0"], p["h1"], p["chol"], rdm1, dt)
    return sys, ham, tr, mops, tops, pops, params, pctx, mctx, ref_ops, ref_ctx


def _walkers(p, kind):
    if kind == "unrestricted":
        return (p["wa"], p["wb"])
    return p["wa"]


def _to_dev(w):
    if isinstance(w, tuple):
        return tuple(torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in w)
    return torch.from_numpy(np.ascontiguousarray(w)).cuda()


def _cmp_walkers(wg, wr):
    if isinstance(wr, tuple):
        return max(rel(wg[0].cpu().numpy(), wr[0]), rel(wg[1].cpu().numpy(), wr[1]))
    return rel(wg.cpu().numpy(), wr)


@pytest.mark.parametrize("layout", ["auto", "dense"])
@pytest.mark.parametrize("case", [
    ("restricted", (24, 6, 40, 16), False),
    ("restricted", (37, 5, 23, 70), False),
    ("unrestricted", (30, 5, 28, 9), False),
    ("unrestricted", (33, 7, 30, 11), True),
    ("nonsym", (21, 4, 13, 10), False),
])
def test_steps_match_oracle(gpu, case, layout):
    kind, (N, n, G, W), uhf = case
    p = problem(N, n, G, W, seed=21, nocc_b=(n - 1 if uhf else (n if kind == "unrestricted" else None)))
    if kind == "nonsym":  # non-symmetric L_g (trot tests/test_prop_afqmc_step.py:85-162): dense VHS path
        kind = "restricted"
        p["chol"] = p["chol"] + 0.01 * np.random.default_rng(5).standard_normal(p["chol"].shape)
    sys, ham, tr, mops, tops, pops, params, pctx, mctx, ref_ops, ref_ctx = _setup(p, kind, 0.01, uhf, layout)
    assert ("chol_symblk" in pctx.low) == (layout == "auto" and bool(np.array_equal(
        p["chol"], np.swapaxes(p["chol"], 1, 2))))
    w0 = _walkers(p, kind)
    state = pops.init_prop_state(sys=sys, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                 params=params, initial_walkers=_to_dev(w0))
    ref = O.init_prop_state(w0, ref_ops, W)
    assert abs(state.e_estimate.item() - ref.e_estimate) < 1e-10 * abs(ref.e_estimate)
    ref_params = O.QmcParams(dt=0.01, n_walkers=W)
    for s in range(6):
        f = p["fields"][s]
        state = pops.step(state, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                          meas_ctx=mctx, prop_ctx=pctx, fields=torch.from_numpy(f).cuda())
        ref = O.afqmc_step(ref, f, ref_ctx, ref_ops, ref_params)
        assert _cmp_walkers(state.walkers, ref.walkers) < TOL, s
        assert rel_each(state.overlaps.cpu().numpy(), ref.overlaps) < TOL, s
        assert rel_each(state.weights.cpu().numpy(), ref.weights) < TOL, s
        assert abs(state.pop_control_ene_shift.item() - ref.pop_control_ene_shift) < 1e-10 * max(
            1.0, abs(ref.pop_control_ene_shift)), s
        assert int(state.node_encounters.item()) == ref.node_encounters


def test_step_does_not_modify_input_state(gpu):
    p = problem(20, 4, 15, 6, seed=22)
    sys, ham, tr, mops, tops, pops, params, pctx, mctx, _, _ = _setup(p, "restricted", 0.01)
    state = pops.init_prop_state(sys=sys, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                 params=params, initial_walkers=_to_dev(p["wa"]))
    snap = [state.walkers.clone(), state.weights.clone(), state.overlaps.clone()]
    f = torch.from_numpy(p["fields"][0]).cuda()
    out1 = pops.step(state, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                     meas_ctx=mctx, prop_ctx=pctx, fields=f)
    out2 = pops.step(state, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                     meas_ctx=mctx, prop_ctx=pctx, fields=f)
    assert torch.equal(snap[0], state.walkers) and torch.equal(snap[1], state.weights)
    assert torch.equal(snap[2], state.overlaps)
    # deterministic: the same inputs give bit-identical outputs
    assert torch.equal(out1.walkers, out2.walkers) and torch.equal(out1.weights, out2.weights)


def test_zero_hamiltonian_known_answer(gpu):
    """trot tests/test_prop_afqmc_step.py:33-82: chol = 0, h0 = 1, dt = 0.2 -> w = e^-0.2, shift = 0.1."""
    from trot_amd import meas, prop, trial
    from trot_amd.core import HamChol, System

    N, n, G, W = 4, 2, 3, 8
    ham = HamChol(1.0, np.zeros((N, N)), np.zeros((G, N, N))).to_device()
    sys = System(N, (n, n), "restricted")
    c = np.eye(N)[:, :n]
    tr = trial.RhfTrial(c)
    mops = meas.make_rhf_meas_ops(sys)
    tops = trial.make_rhf_trial_ops(sys)
    pops = prop.make_prop_ops("restricted", "restricted")
    params = prop.QmcParams(dt=0.2, n_exp_terms=4, pop_control_damping=0.1, n_walkers=W, seed=0)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    w = torch.from_numpy(np.broadcast_to(c + 0j, (W, N, n)).copy()).cuda()
    state = pops.init_prop_state(sys=sys, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                 params=params, initial_walkers=w, initial_e_estimate=0.0)
    out = pops.step(state, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                    meas_ctx=mctx, prop_ctx=pctx)
    assert np.allclose(out.weights.cpu().numpy(), np.exp(-0.2), rtol=1e-14)
    assert abs(out.pop_control_ene_shift.item() - 0.1) < 1e-14
    assert out.e_estimate.item() == 0.0
    assert abs(mops.require_kernel("energy")(w, ham, mctx, tr).cpu().numpy() - 1.0).max() < 1e-14


@pytest.mark.parametrize("kind", ["restricted", "unrestricted"])
def test_block_matches_oracle(gpu, kind):
    from trot_amd import blocks

    N, n, G, W = 26, 5, 30, 24
    p = problem(N, n, G, W, seed=23, nocc_b=(n if kind == "unrestricted" else None))
    sys, ham, tr, mops, tops, pops, params, pctx, mctx, ref_ops, ref_ctx = _setup(p, kind, 0.01)
    w0 = _walkers(p, kind)
    state = pops.init_prop_state(sys=sys, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                 params=params, initial_walkers=_to_dev(w0))
    ref = O.init_prop_state(w0, ref_ops, W)
    fseq = p["fields"][:params.n_prop_steps]
    seed, ctr = state.rng_key
    zeta = float(pctx.fields.uniform((seed, ctr + params.n_prop_steps))[0])  # the block's SR draw
    state, obs = blocks.block(state, sys=sys, params=params, ham_data=ham, trial_data=tr, trial_ops=tops,
                              meas_ops=mops, meas_ctx=mctx, prop_ops=pops, prop_ctx=pctx,
                              fields_seq=[torch.from_numpy(f).cuda() for f in fseq])
    ref_params = O.QmcParams(dt=0.01, n_walkers=W, n_prop_steps=params.n_prop_steps)
    ref, robs = O.block(ref, fseq, zeta, ref_ctx, ref_ops, ref_params)
    assert abs(obs.scalars["energy"].item() - robs["energy"]) < 1e-10 * abs(robs["energy"])
    assert abs(obs.scalars["weight"].item() - robs["weight"]) < 1e-10 * abs(robs["weight"])
    assert abs(state.e_estimate.item() - ref.e_estimate) < 1e-10 * abs(ref.e_estimate)
    assert _cmp_walkers(state.walkers, ref.walkers) < TOL
    assert rel_each(state.overlaps.cpu().numpy(), ref.overlaps) < TOL
    assert rel_each(state.weights.cpu().numpy(), ref.weights) < 1e-13


def test_run_qmc_short(gpu):
    """The driver runs equilibration + sampling blocks and returns finite statistics."""
    from trot_amd import blocks, driver

    p = problem(16, 4, 20, 32, seed=24)
    sys, ham, tr, mops, tops, pops, params, pctx, mctx, _, _ = _setup(p, "restricted", 0.01)
    from trot_amd.prop import QmcParams

    params = QmcParams(dt=0.01, n_walkers=32, seed=5, n_prop_steps=3, n_eql_blocks=2, n_blocks=30)
    res = driver.run_qmc(sys=sys, params=params, ham_data=ham, trial_data=tr, meas_ops=mops, trial_ops=tops,
                         prop_ops=pops, block_fn=blocks.block, verbose=False)
    assert np.isfinite(res.mean_energy)
    assert res.block_energies.shape[0] >= 20
