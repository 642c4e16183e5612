"""The molecular parity case: linear H12, bond 1.8 bohr, STO-6G (12 orbitals, 6 occupied per spin).

Hamiltonian, trial and Cholesky vectors (cut 1e-5) are staged by oracle/h_chain.py, which restates
trot/staging.py's RHF path.  64 complex128 walkers start on the trial, dt = 0.005, 50 steps per block,
20 blocks, N(0,1) fields from numpy's default_rng(0) drawn step by step.  The FP64 GPU path and the
CPU oracle take identical fields and SR draws; trot orthonormalizes once per block (blocks.py:230).
Checked: walkers, weights and E_L at 1e-10 relative after the first block, and every block energy,
the weights and the walkers after all 20 blocks.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import afqmc_oracle as O
from oracle import h_chain as H
from tests.cases import rel, rel_each

pytestmark = pytest.mark.gpu

W, DT, NPS, NBLK = 64, 0.005, 50, 20


@pytest.fixture(scope="module")
def h12():
    return H.stage_rhf(H.chain(12, 1.8), 12, chol_cut=1e-5)


def test_h12_blocks_match_oracle(gpu, h12):
    from trot_amd import blocks, meas, prop, trial
    from trot_amd.core import HamChol, System

    st = h12
    N, n = st.h1.shape[0], st.nocc
    G = st.chol.shape[0]
    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(st.h0, st.h1, st.chol).to_device()
    tr = trial.RhfTrial(st.trial)
    mops = meas.make_rhf_meas_ops(sys_)
    tops = trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted")
    params = prop.QmcParams(dt=DT, n_walkers=W, seed=0, n_prop_steps=NPS)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    assert pctx.low.get("symblk")  # real-orbital ERIs: symmetric Cholesky vectors

    ref_ops = O.make_rhf_ops(st.h0, st.h1, st.chol, st.trial)
    ref_ctx = O.build_prop_ctx(st.h0, st.h1, st.chol, O.rhf_rdm1(st.trial), DT)
    ref_params = O.QmcParams(dt=DT, n_walkers=W, n_prop_steps=NPS)
    w0 = np.repeat(st.trial[None].astype(np.complex128), W, axis=0)
    state = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                 params=params, initial_walkers=torch.from_numpy(w0).cuda())
    ref = O.init_prop_state(w0, ref_ops, W)
    rng = np.random.default_rng(0)
    e_gpu, e_ref = [], []
    for b in range(NBLK):
        fseq = [rng.standard_normal((W, G)) for _ in range(NPS)]
        seed, ctr = state.rng_key
        zeta = float(pctx.fields.uniform((seed, ctr + NPS))[0])  # the draw the block's SR uses
        state, obs = blocks.block(state, sys=sys_, params=params, ham_data=ham, trial_data=tr, trial_ops=tops,
                                  meas_ops=mops, meas_ctx=mctx, prop_ops=pops, prop_ctx=pctx,
                                  fields_seq=[torch.from_numpy(f).cuda() for f in fseq])
        ref, robs = O.block(ref, fseq, zeta, ref_ctx, ref_ops, ref_params)
        e_gpu.append(obs.scalars["energy"].item())
        e_ref.append(robs["energy"])
        if b == 0:
            assert rel(state.walkers.cpu().numpy(), ref.walkers) < 1e-10
            assert rel_each(state.weights.cpu().numpy(), ref.weights) < 1e-10
            e_l = mops.require_kernel("energy")(state.walkers, ham, mctx, tr).cpu().numpy()
            assert rel(e_l, ref_ops.energy(ref.walkers)) < 1e-10
    e_gpu, e_ref = np.array(e_gpu), np.array(e_ref)
    print(f"[h12] E_RHF {st.e_rhf:.8f}  nchol {G}  block energies gpu {e_gpu.mean():.8f} oracle {e_ref.mean():.8f}"
          f"  max |dE| {np.abs(e_gpu - e_ref).max():.3e}", flush=True)
    assert np.all(np.abs(e_gpu - e_ref) < 1e-10 * np.abs(e_ref))
    assert rel_each(state.weights.cpu().numpy(), ref.weights) < 1e-9
    assert rel(state.walkers.cpu().numpy(), ref.walkers) < 1e-9
    # the propagated estimate sits below the mean-field energy, as a correlated estimate should
    assert e_ref[NBLK // 2:].mean() < st.e_rhf
