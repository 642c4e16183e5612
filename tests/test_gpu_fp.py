"""Free projection on the GPU against the oracle (FP64 bar 1e-10), with pre-drawn fields and comb offsets.

* one step (restricted walkers with an RHF trial, unrestricted walkers with a UHF trial): walkers
  after the comb, weights with the norm divided back out;
* a block of 4 steps in 2 QR groups: walkers, weights, overlaps and the four block sums
  (energy, weight, overlap, |overlap|);
* the kernels that are new here, one at a time: the constants and their root, the QR norm, the
  orthogonalization norm;
* run_qmc_fp on the device stream (2 trajectories x 2 blocks): finite, the shapes of the reference's
  result, trajectory 0's first block equal to a direct block_fp call from the same state.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import afqmc_oracle as O
from tests.cases import problem, rel, rel_each

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _setup(kind, N, n, G, W, ene0=-1.0, dt=0.01, nps=4, nqr=2):
    from trot_amd import fp, meas, trial
    from trot_amd.core import HamChol, System

    uhf = kind == "unrestricted"
    p = problem(N, n, G, W, seed=51, nocc_b=(n - 1 if uhf else None))
    nb = n - 1 if uhf else n
    sys = System(N, (n, nb), kind)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    if uhf:
        tr = trial.UhfTrial(p["ca"], p["cb"])
        mops = meas.make_uhf_meas_ops(sys)
        tops = trial.make_uhf_trial_ops(sys)
        ref_ops = O.make_uhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], p["cb"])
        rdm1 = O.uhf_rdm1(p["ca"], p["cb"])
        w0 = (p["wa"], p["wb"])
    else:
        tr = trial.RhfTrial(p["ca"])
        mops = meas.make_rhf_meas_ops(sys)
        tops = trial.make_rhf_trial_ops(sys)
        ref_ops = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"])
        rdm1 = O.rhf_rdm1(p["ca"])
        w0 = p["wa"]
    params = fp.QmcParamsFp(dt=dt, n_walkers=W, seed=5, n_prop_steps=nps, n_qr_blocks=nqr, n_blocks=2, n_traj=2,
                            ene0=ene0)
    pops = fp.make_prop_ops_fp("restricted", kind, sys)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    ref_ctx = O.build_prop_ctx_fp(p["h0"], p["h1"], p["chol"], rdm1, dt, ene0)
    rng = np.random.default_rng(7)
    fields = [rng.standard_normal((W, G)) for _ in range(nps)]
    zetas = list(rng.uniform(size=nps))
    return dict(p=p, sys=sys, ham=ham, tr=tr, mops=mops, tops=tops, pops=pops, params=params, pctx=pctx, mctx=mctx,
                ref_ops=ref_ops, ref_ctx=ref_ctx, w0=w0, fields=fields, zetas=zetas)


def _dev(w):
    return tuple(_t(x) for x in w) if isinstance(w, tuple) else _t(w)


def _cmp(wg, wr):
    if isinstance(wr, tuple):
        return max(rel(wg[0].cpu().numpy(), wr[0]), rel(wg[1].cpu().numpy(), wr[1]))
    return rel(wg.cpu().numpy(), wr)


def _states(c):
    st = c["pops"].init_prop_state(sys=c["sys"], ham_data=c["ham"], trial_ops=c["tops"], trial_data=c["tr"],
                                   meas_ops=c["mops"], params=c["params"], initial_walkers=_dev(c["w0"]))
    ref = O.init_prop_state(c["w0"], c["ref_ops"], c["params"].n_walkers)
    return st, ref


@pytest.mark.parametrize("kind,shape", [("restricted", (24, 5, 30, 12)), ("restricted", (37, 6, 23, 9)),
                                        ("unrestricted", (26, 5, 28, 10))])
def test_fp_step_matches_oracle(gpu, kind, shape):
    print(f"[fp step] {kind} N,n,G,W={shape}", flush=True)
    c = _setup(kind, *shape)
    st, ref = _states(c)
    for s in range(2):
        st = c["pops"].step(st, params=c["params"], ham_data=c["ham"], trial_data=c["tr"], trial_ops=c["tops"],
                            meas_ops=c["mops"], meas_ctx=c["mctx"], prop_ctx=c["pctx"], fields=_t(c["fields"][s]),
                            zeta=float(c["zetas"][s]))
        ref = O.afqmc_step_fp(ref, c["fields"][s], float(c["zetas"][s]), c["ref_ctx"], kind)
        assert _cmp(st.walkers, ref.walkers) < TOL, s
        assert rel_each(st.weights.cpu().numpy(), ref.weights) < TOL, s
    assert st.rng_key[1] == 4  # one field draw and one comb offset per step


@pytest.mark.parametrize("kind", ["restricted", "unrestricted"])
def test_fp_block_matches_oracle(gpu, kind):
    from trot_amd import fp

    print(f"[fp block] {kind}", flush=True)
    c = _setup(kind, 24, 5, 30, 12)
    st, ref = _states(c)
    st, obs = fp.block_fp(st, sys=c["sys"], params=c["params"], ham_data=c["ham"], trial_data=c["tr"],
                          trial_ops=c["tops"], meas_ops=c["mops"], meas_ctx=c["mctx"], prop_ops=c["pops"],
                          prop_ctx=c["pctx"], fields_seq=[_t(f) for f in c["fields"]],
                          zetas=[float(z) for z in c["zetas"]])
    ref_params = O.QmcParamsFp(dt=0.01, n_prop_steps=4, n_qr_blocks=2, ene0=-1.0, n_walkers=12)
    ref, robs = O.block_fp(ref, c["fields"], [float(z) for z in c["zetas"]], c["ref_ctx"], c["ref_ops"], ref_params)
    assert _cmp(st.walkers, ref.walkers) < TOL
    assert rel_each(st.weights.cpu().numpy(), ref.weights) < TOL
    assert rel_each(st.overlaps.cpu().numpy(), ref.overlaps) < TOL
    for k in ("energy", "weight", "overlap", "abs_overlap"):
        g = complex(obs.scalars[k].item())
        assert abs(g - complex(robs[k])) <= TOL * max(abs(complex(robs[k])), 1e-300), k


def test_fp_kernels_one_at_a_time(gpu):
    from trot_amd import _lib

    rng = np.random.default_rng(9)
    W, G, N, n = 7, 33, 11, 4
    x = rng.standard_normal((W, G))
    mf = 1j * rng.standard_normal(G) * 0.1
    sdt, dt, h0p = 0.1, 0.01, complex(-0.3, 0.05)
    xbar, cst, scale = _lib.fp_prepare(_t(x), _t(mf), sdt, dt, h0p, 2 * n)
    c_ref = np.exp(-sdt * (x @ mf) + dt * h0p)
    assert rel(cst.cpu().numpy(), c_ref) < 1e-13
    assert rel(scale.cpu().numpy(), c_ref ** (1.0 / (2 * n))) < 1e-13
    assert np.array_equal(xbar.cpu().numpy(), x + 0j)
    w = rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n))
    wt = _t(w)
    _lib.scale_rows(wt, scale)
    assert rel(wt.cpu().numpy(), w * (c_ref ** (1.0 / (2 * n)))[:, None, None]) < 1e-13
    d = _lib.qr_norm(_t(w)).cpu().numpy()
    assert rel(d * d, O.qr_norm(w, "restricted").real) < 1e-12
    q, d2 = _lib.orthogonalize(_t(w))
    qr, nr = O.orthogonalize(w, "restricted")
    assert rel(q.cpu().numpy(), qr) < 1e-12
    assert rel(d2.cpu().numpy() ** 2, nr.real) < 1e-12


def test_run_qmc_fp(gpu):
    from trot_amd import fp

    c = _setup("restricted", 24, 5, 30, 12)
    args = dict(sys=c["sys"], params=c["params"], ham_data=c["ham"], trial_data=c["tr"], meas_ops=c["mops"],
                trial_ops=c["tops"], prop_ops=c["pops"], meas_ctx=c["mctx"], prop_ctx=c["pctx"])
    st0, _ = _states(c)
    res = fp.run_qmc_fp(state=st0, verbose=False, **args)
    assert res.block_energies.shape == (2, 3) and res.mean_energy.shape == (3,)
    assert np.all(np.isfinite(res.block_energies)) and np.all(np.isfinite(res.mean_energy))
    st0, _ = _states(c)
    _, obs = fp.block_fp(st0, **{k: v for k, v in args.items()})
    assert complex(obs.scalars["energy"].item()) == res.block_energies[0, 1]
