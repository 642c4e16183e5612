"""The differentiation form: dE_block/dlambda for h1 + lambda D by tangent propagation (trot_amd/diff.py).

Sizes of the differentiation case: N=128, n=16, G=640, 512 walkers, dt = 0.005, 200 steps, synthetic
system of seed 3 (tau 256) with every L_g scaled by 1/sqrt(N) (the recipe as written drives every
walker's importance out of [floor, cap] within ~10 steps, so the block energy would be constant),
D symmetric N(0,1) * 0.05.  Checked against central finite differences at lambda = +-1e-4 with common
fields, computed on the product's own step and block path (not the tangent code): relative 1e-4.  The
tangent run's own block energy at lambda = 0 equals that path's to 1e-10; resuming from the checkpoint at
step 100 reproduces the result bit for bit; the autograd Function returns the same derivative.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N, n, G, W, DT, STEPS, H = 128, 16, 640, 512, 0.005, 200, 1e-4


def _system():
    from trot_amd import synthetic

    syn = synthetic.make_synthetic(N, n, G, W, 3, 256.0, chol_scale=1.0 / np.sqrt(N))
    rng = np.random.default_rng(33)
    a = rng.standard_normal((N, N))
    D = 0.05 * (np.triu(a) + np.triu(a, 1).T)
    return syn, D


def _fields(s):
    from trot_amd import synthetic

    return torch.from_numpy(synthetic.fields(3, s, W, G)).cuda()


def _product_block_energy(syn, D, lam):
    from trot_amd import blocks, meas, prop, trial
    from trot_amd.core import HamChol, System

    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(syn.h0, syn.h1 + lam * D, syn.chol).to_device()
    tr = trial.RhfTrial(syn.trial)
    mops = meas.make_rhf_meas_ops(sys_)
    tops = trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted")
    params = prop.QmcParams(dt=DT, n_walkers=W, seed=3, n_prop_steps=STEPS)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                              params=params, initial_walkers=torch.from_numpy(syn.walkers).cuda())
    st, obs = blocks.block(st, sys=sys_, params=params, ham_data=ham, trial_data=tr, trial_ops=tops,
                           meas_ops=mops, meas_ctx=mctx, prop_ops=pops, prop_ctx=pctx,
                           fields_seq=[_fields(s) for s in range(STEPS)])
    return float(obs.scalars["energy"].item())


def test_block_energy_derivative(gpu):
    from trot_amd import diff, prop

    syn, D = _system()
    params = prop.QmcParams(dt=DT, n_walkers=W, seed=3, n_prop_steps=STEPS)
    w0 = torch.from_numpy(syn.walkers).cuda()

    def run_at(lam, **kw):
        tc = diff.build_ctx(syn.h0, syn.h1, syn.chol, syn.trial, D, lam, params)
        return tc, diff.run(tc, w0, _fields, STEPS, **kw)

    tc, res = run_at(0.0, checkpoint_every=20)
    e0 = _product_block_energy(syn, D, 0.0)
    ep = _product_block_energy(syn, D, H)
    em = _product_block_energy(syn, D, -H)
    fd = (ep - em) / (2 * H)
    print(f"[diff] E_block {res.energy:.10f} (product path {e0:.10f})  dE/dlambda tangent {res.denergy:.8e}  "
          f"central difference {fd:.8e}  rel {abs(res.denergy - fd) / abs(fd):.2e}", flush=True)
    assert abs(res.energy - e0) <= 1e-10 * abs(e0)
    assert abs(res.denergy - fd) <= 1e-4 * abs(fd)
    assert len(res.checkpoints) == STEPS // 20
    again = diff.run(tc, None, _fields, STEPS, resume=res.checkpoints[4])  # from step 100
    assert again.energy == res.energy and again.denergy == res.denergy

    lam = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    e = diff.BlockEnergy.apply(lam, lambda x: res if x == 0.0 else run_at(x)[1])
    e.backward()
    assert e.item() == res.energy and lam.grad.item() == res.denergy
