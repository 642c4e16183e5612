"""The CISD trial on the GPU (trot_amd/cisd.py, trot_amd/csrc/cisd.hip) against the CISD oracle.

FP64 bar 1e-10 relative for the overlap, force bias, local energy and mixed 1-RDM, with and without
frozen core / outer orbitals, with the energy's walker chunks smaller than the batch; the split-BF16
force bias and doubles term (mixed_precision=True, testing=False: the reference's defaults) within the
tolerances the reference allows its float32 variants (tests/test_cisd.py:245-287, :426); one propagation step pair and one block (energy, QR,
SR) through the generic step and block driver against the oracle's.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import afqmc_oracle as O
from oracle import cisd_oracle as C
from tests.cases import rel

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _case(N, nocc, core, outer, G, W, seed, s1=0.1, s2=0.05, mix=0.2):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((N, N)) / np.sqrt(N)
    h1 = 0.5 * (a + a.T)
    b = rng.standard_normal((G, N, N))
    chol = 0.5 * (b + np.swapaxes(b, 1, 2)) * 0.1 * np.exp(-np.arange(G) / G)[:, None, None]
    na, va = nocc - core, N - nocc - outer
    ci1 = s1 * rng.standard_normal((na, va))
    ci2 = s2 * rng.standard_normal((na, va, na, va))
    ci2 = 0.5 * (ci2 + ci2.transpose(2, 3, 0, 1))
    w = np.zeros((W, N, nocc), complex)
    w[:, :nocc] = np.eye(nocc)
    w = w + mix * (rng.standard_normal((W, N, nocc)) + 1j * rng.standard_normal((W, N, nocc)))
    t = C.CisdTrial(ci1, ci2, core, outer)
    return dict(h0=0.37, h1=h1, chol=chol, trial=t, w=w, rng=rng)


def _gpu(c, mixed=False, testing=None):
    from trot_amd import cisd
    from trot_amd.core import HamChol, System

    t = c["trial"]
    N, n = t.norb, t.nocc_full
    sys = System(N, (n, n), "restricted")
    ham = HamChol(c["h0"], c["h1"], c["chol"]).to_device()
    tr = cisd.make_cisd_trial_data(dict(ci1=t.ci1, ci2=t.ci2, nocc_t_core=t.nocc_t_core,
                                        nvir_t_outer=t.nvir_t_outer), sys)
    mops = cisd.make_cisd_meas_ops(sys, mixed_precision=mixed, testing=(not mixed) if testing is None else testing)
    tops = cisd.make_cisd_trial_ops(sys)
    return sys, ham, tr, mops, tops, mops.build_meas_ctx(ham, tr)


SHAPES = [(12, 4, 0, 0, 10, 16), (20, 6, 1, 3, 24, 40), (37, 9, 2, 5, 45, 70)]


@pytest.mark.parametrize("shape", SHAPES)
def test_measurement_matches_oracle(gpu, shape):
    from trot_amd import cisd
    from trot_amd.core import k_energy, k_force_bias, o_rdm1

    print(f"[cisd meas] N,nocc,core,outer,G,W={shape}", flush=True)
    c = _case(*shape, seed=sum(shape))
    sys, ham, tr, mops, tops, mctx = _gpu(c)
    t, w = c["trial"], c["w"]
    wd = torch.from_numpy(w).cuda()
    ctx = C.build_meas_ctx(c["chol"], t)
    assert rel(tops.overlap(wd, tr).cpu().numpy(), C.overlap(w, t)) < TOL
    assert rel(mops.require_kernel(k_force_bias)(wd, ham, mctx, tr).cpu().numpy(),
               C.force_bias(w, c["chol"], ctx, t)) < TOL
    e_ref = C.energy(w, c["h0"], c["h1"], c["chol"], ctx, t)
    assert rel(mops.require_kernel(k_energy)(wd, ham, mctx, tr).cpu().numpy(), e_ref) < TOL
    assert rel(mops.observables[o_rdm1](wd, ham, mctx, tr).cpu().numpy(), C.rdm1(w, t)) < TOL
    # one walker (per-walker kernel form) and walker chunks of 7 in the energy
    assert abs(complex(cisd.energy_kernel_rw_rh(wd[3], ham, mctx, tr).item()) - e_ref[3]) < TOL * abs(e_ref[3])
    eng = mops.engine
    _, greens = eng.overlap_green(wd, tr)
    gt, bs, om = greens
    from trot_amd import _lib

    e7 = _lib.cisd_energy(gt, bs, _lib.cisd_ps(gt, bs, t.nocc_t_core), om, mctx.lx, tr.dmat(), t.nocc_t_core,
                          c["h0"], wc=7)
    assert rel(e7.cpu().numpy(), e_ref) < TOL


def test_mixed_precision_force_bias_and_energy(gpu):
    """The reference's defaults (mixed_precision=True, testing=False): split-BF16 force bias within the 2e-5
    its float32 variant is held to; the energy with the doubles term on split-BF16 within the rtol 2e-5 /
    atol 2e-6 the reference holds its default (float32) energy to."""
    from trot_amd.core import k_energy, k_force_bias

    c = _case(30, 6, 1, 2, 40, 24, seed=77)
    sys, ham, tr, mops, tops, mctx = _gpu(c, mixed=True)
    assert mops.engine.precision == "bf16x3" and mops.engine.doubles_precision == "bf16x3"
    wd = torch.from_numpy(c["w"]).cuda()
    ctx = C.build_meas_ctx(c["chol"], c["trial"])
    fb = mops.require_kernel(k_force_bias)(wd, ham, mctx, tr).cpu().numpy()
    ref = C.force_bias(c["w"], c["chol"], ctx, c["trial"])
    assert np.allclose(fb, ref, rtol=2e-5, atol=2e-5)
    assert rel(fb, ref) > 0  # the split-BF16 path ran, not the FP64 one
    e = mops.require_kernel(k_energy)(wd, ham, mctx, tr).cpu().numpy()
    e_ref = C.energy(c["w"], c["h0"], c["h1"], c["chol"], ctx, c["trial"])
    assert rel(e, e_ref) > 0 and np.allclose(e, e_ref, rtol=2e-5, atol=2e-6)  # tests/test_cisd.py:426


def _prop(c, sys, ham, tr, mops, tops, dt=0.01, n_prop_steps=3):
    from trot_amd import prop

    W = c["w"].shape[0]
    pops = prop.make_prop_ops("restricted", "restricted")
    params = prop.QmcParams(dt=dt, n_walkers=W, seed=3, n_prop_steps=n_prop_steps)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    ref_ctx = O.build_prop_ctx(c["h0"], c["h1"], c["chol"], C.get_rdm1(c["trial"]), dt)
    ref_ops = C.make_ops(c["h0"], c["h1"], c["chol"], c["trial"])
    return pops, params, pctx, ref_ctx, ref_ops


def test_steps_match_oracle(gpu):
    c = _case(24, 6, 1, 2, 30, 32, seed=5)
    sys, ham, tr, mops, tops, mctx = _gpu(c)
    pops, params, pctx, ref_ctx, ref_ops = _prop(c, sys, ham, tr, mops, tops)
    W = c["w"].shape[0]
    st = pops.init_prop_state(sys=sys, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=torch.from_numpy(c["w"]).cuda())
    ref = O.init_prop_state(c["w"], ref_ops, W)
    assert abs(st.e_estimate.item() - ref.e_estimate) < TOL * abs(ref.e_estimate)
    ref_params = O.QmcParams(dt=0.01, n_walkers=W)
    for s in range(2):
        f = c["rng"].standard_normal((W, 30))
        st = pops.step(st, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
                       prop_ctx=pctx, fields=torch.from_numpy(f).cuda())
        ref = O.afqmc_step(ref, f, ref_ctx, ref_ops, ref_params)
        assert rel(st.walkers.cpu().numpy(), ref.walkers) < TOL, s
        assert rel(st.overlaps.cpu().numpy(), ref.overlaps) < TOL, s
        assert rel(st.weights.cpu().numpy(), ref.weights) < TOL, s
        assert abs(st.pop_control_ene_shift.item() - ref.pop_control_ene_shift) < TOL * abs(ref.pop_control_ene_shift)


def test_block_matches_oracle(gpu):
    from trot_amd import blocks

    c = _case(20, 5, 0, 1, 26, 24, seed=9)
    sys, ham, tr, mops, tops, mctx = _gpu(c)
    pops, params, pctx, ref_ctx, ref_ops = _prop(c, sys, ham, tr, mops, tops)
    W = c["w"].shape[0]
    st = pops.init_prop_state(sys=sys, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=torch.from_numpy(c["w"]).cuda())
    ref = O.init_prop_state(c["w"], ref_ops, W)
    fseq = [c["rng"].standard_normal((W, 26)) for _ in range(params.n_prop_steps)]
    seed, ctr = st.rng_key
    zeta = float(pctx.fields.uniform((seed, ctr + params.n_prop_steps))[0])
    st, obs = blocks.block(st, sys=sys, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                           meas_ctx=mctx, prop_ops=pops, prop_ctx=pctx,
                           fields_seq=[torch.from_numpy(f).cuda() for f in fseq])
    ref, robs = O.block(ref, fseq, zeta, ref_ctx, ref_ops, O.QmcParams(dt=0.01, n_walkers=W, n_prop_steps=3))
    assert abs(obs.scalars["energy"].item() - robs["energy"]) < TOL * abs(robs["energy"])
    assert abs(obs.scalars["weight"].item() - robs["weight"]) < TOL * abs(robs["weight"])
    assert rel(st.walkers.cpu().numpy(), ref.walkers) < TOL
    assert rel(st.overlaps.cpu().numpy(), ref.overlaps) < TOL
    assert rel(st.weights.cpu().numpy(), ref.weights) < 1e-13
