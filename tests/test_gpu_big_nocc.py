"""Occupied spaces beyond one 64-wide block: nocc = 65 .. 200 (the reference handles any nocc through
jnp.linalg.solve and det, trot/meas/rhf.py:21-25, trot/trial/rhf.py:35-47).

* the inverse and determinant (LDS-resident Gauss-Jordan up to n = 96, in place in the output above) against
  numpy, including matrices that need row interchanges;
* overlaps, force bias and local energy (FP64; the exchange in 64 x 64 tile pairs) against the oracle for RHF
  with restricted and unrestricted walkers and for a UHF trial with na != nb on both sides of 64;
* the QR (Gram-Schmidt with reorthogonalisation above 64) and |det R| against the oracle's qr_phase;
* two FP64 steps and one block (QR, energy, SR) at nocc = 96 against the oracle at 1e-10;
* the split-BF16 and FP32 energies (FP32-MFMA exchange tiles above 64) within their stated bounds, and two
  split-BF16 steps within the propagation bound.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import afqmc_oracle as O
from tests.cases import problem, rel, rel_each
from tests.test_gpu_step import _setup, _to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("n", [65, 96, 97, 128, 200])
def test_inverse_det_beyond_64(gpu, n):
    from trot_amd import _lib

    rng = np.random.default_rng(n)
    W = 5
    m = rng.standard_normal((W, n, n)) + 1j * rng.standard_normal((W, n, n))
    m[1] = m[1][rng.permutation(n)] * 0.5  # row interchanges everywhere, scaled entries
    m[2, :, 0] *= 1e-9  # a tiny first column: the pivot search must swap at every step's start
    minv, det = _lib.inverse_det(torch.from_numpy(m).cuda())
    minv, det = minv.cpu().numpy(), det.cpu().numpy()
    for b in range(W):
        assert rel(minv[b], np.linalg.inv(m[b])) < 1e-10, b
        assert abs(det[b] - np.linalg.det(m[b])) < 1e-10 * abs(np.linalg.det(m[b])), b


@pytest.mark.parametrize("kind,n,nb", [("restricted", 65, None), ("restricted", 96, None), ("restricted", 128, None),
                                       ("unrestricted", 96, 96), ("uhf", 96, 70), ("uhf", 70, 40)])
def test_overlap_force_bias_energy_beyond_64(gpu, kind, n, nb):
    from trot_amd import meas, trial
    from trot_amd.core import HamChol, System

    N, G, W = 160, 40, 6
    p = problem(N, n, G, W, seed=50 + n, nocc_b=nb)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    if kind == "uhf":
        sys_ = System(N, (n, nb), "unrestricted")
        tr = trial.UhfTrial(p["ca"], p["cb"])
        ops = meas.make_uhf_meas_ops(sys_)
        ref = O.make_uhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], p["cb"])
        w, wn = (_to_dev(p["wa"]), _to_dev(p["wb"])), (p["wa"], p["wb"])
    else:
        sys_ = System(N, (n, n), kind)
        tr = trial.RhfTrial(p["ca"])
        ops = meas.make_rhf_meas_ops(sys_)
        ref = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], kind)
        w, wn = ((_to_dev(p["wa"]), _to_dev(p["wb"])), (p["wa"], p["wb"])) if kind == "unrestricted" else \
            (_to_dev(p["wa"]), p["wa"])
    mctx = ops.build_meas_ctx(ham, tr)
    assert rel_each(ops.overlap(w, tr).cpu().numpy(), ref.overlap(wn)) < TOL
    assert rel(ops.require_kernel("force_bias")(w, ham, mctx, tr).cpu().numpy(), ref.force_bias(wn)) < TOL
    e = ops.require_kernel("energy")(w, ham, mctx, tr).cpu().numpy()
    assert rel(e - p["h0"], ref.energy(wn) - p["h0"]) < TOL


@pytest.mark.parametrize("n", [65, 100])
def test_qr_beyond_64(gpu, n):
    from trot_amd import _lib

    rng = np.random.default_rng(n)
    W, N = 4, 150
    w = rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n))
    q = _lib.orthonormalize(_to_dev(w)).cpu().numpy()
    q_ref, det_ref = O.qr_phase(w)
    assert rel(q, q_ref) < TOL
    d = _lib.qr_norm(_to_dev(w)).cpu().numpy()
    assert rel_each(d, np.abs(det_ref)) < TOL
    q2, d2 = _lib.orthogonalize(_to_dev(w))
    assert rel(q2.cpu().numpy(), q_ref) < TOL and rel_each(d2.cpu().numpy(), np.abs(det_ref)) < TOL


def test_steps_and_block_at_nocc_96(gpu):
    from trot_amd import blocks

    N, n, G, W = 128, 96, 50, 8
    p = problem(N, n, G, W, seed=61)
    sys, ham, tr, mops, tops, pops, params, pctx, mctx, ref_ops, ref_ctx = _setup(p, "restricted", 0.01)
    state = pops.init_prop_state(sys=sys, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                                 initial_walkers=_to_dev(p["wa"]))
    ref = O.init_prop_state(p["wa"], ref_ops, W)
    assert abs(state.e_estimate.item() - ref.e_estimate) < 1e-10 * abs(ref.e_estimate)
    for s in range(2):
        f = p["fields"][s]
        state = pops.step(state, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                          meas_ctx=mctx, prop_ctx=pctx, fields=_to_dev(f))
        ref = O.afqmc_step(ref, f, ref_ctx, ref_ops, O.QmcParams(dt=0.01, n_walkers=W))
        assert rel(state.walkers.cpu().numpy(), ref.walkers) < TOL, s
        assert rel_each(state.weights.cpu().numpy(), ref.weights) < TOL, s
        assert rel_each(state.overlaps.cpu().numpy(), ref.overlaps) < TOL, s
    fseq = p["fields"][2:2 + params.n_prop_steps]
    seed, ctr = state.rng_key
    zeta = float(pctx.fields.uniform((seed, ctr + params.n_prop_steps))[0])
    state, obs = blocks.block(state, sys=sys, params=params, ham_data=ham, trial_data=tr, trial_ops=tops,
                              meas_ops=mops, meas_ctx=mctx, prop_ops=pops, prop_ctx=pctx,
                              fields_seq=[_to_dev(f) for f in fseq])
    ref, robs = O.block(ref, fseq, zeta, ref_ctx, ref_ops,
                        O.QmcParams(dt=0.01, n_walkers=W, n_prop_steps=params.n_prop_steps))
    assert abs(obs.scalars["energy"].item() - robs["energy"]) < 1e-10 * abs(robs["energy"])
    assert rel(state.walkers.cpu().numpy(), ref.walkers) < TOL
    assert rel_each(state.weights.cpu().numpy(), ref.weights) < TOL


def test_reduced_precision_beyond_64(gpu):
    from trot_amd import meas, prop, trial
    from trot_amd.core import HamChol, System

    N, n, G, W = 128, 80, 60, 6
    p = problem(N, n, G, W, seed=62)
    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    ref = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"])
    e_ref = ref.energy(p["wa"])
    for prec, tol in (("f32", 2e-6), ("bf16x3", 2e-4)):
        ops = meas.make_rhf_meas_ops(sys_, precision=prec)
        mctx = ops.build_meas_ctx(ham, tr)
        e = ops.require_kernel("energy")(_to_dev(p["wa"]), ham, mctx, tr).cpu().numpy()
        assert rel(e - p["h0"], e_ref - p["h0"]) < 4 * tol, prec
    # two split-BF16 steps (propagation planes of two 64-column tiles) against the FP64 oracle
    mops = meas.make_rhf_meas_ops(sys_, precision="bf16x3")
    tops = trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted", precision="bf16x3", taylor_precision="bf16x3")
    params = prop.QmcParams(dt=0.01, n_walkers=W, seed=1)
    pctx, mctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params), mops.build_meas_ctx(ham, tr)
    st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=_to_dev(p["wa"]))
    ref_ctx = O.build_prop_ctx(p["h0"], p["h1"], p["chol"], O.rhf_rdm1(p["ca"]), 0.01)
    r = O.init_prop_state(p["wa"], ref, W)
    for s in range(2):
        f = p["fields"][s]
        st = pops.step(st, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
                       prop_ctx=pctx, fields=_to_dev(f))
        r = O.afqmc_step(r, f, ref_ctx, ref, O.QmcParams(dt=0.01, n_walkers=W))
    assert rel(st.walkers.cpu().numpy() - p["wa"], r.walkers - p["wa"]) < 50 * 2e-4
    assert rel_each(st.weights.cpu().numpy(), r.weights) < 50 * 2e-4
