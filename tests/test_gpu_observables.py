"""Block observables on the GPU against the oracle (FP64, relative 1e-10 unless stated).

* The per-walker observable kernels of the RHF / UHF MeasOps ("rdm1", "density_corr";
  trot/meas/rhf.py:191-261, trot/meas/uhf.py:97-160) through ``trot_rdm1_from_green`` and
  ``trot_density_corr``, at sizes off every tile boundary and past one 256-column tile.
* The weighted block mean ``trot_weighted_mean`` (trot/prop/blocks.py:259-268): against NumPy, bit-identical
  on repeated calls, 0 when the weights sum to 0.
* ``blocks.block`` with observable names against the oracle block; ``GraphBlock`` with observables
  against the eager block, bit for bit; ``run_qmc`` collecting block observables and their
  jackknife errors (trot/driver.py:352-400).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import afqmc_oracle as O
from tests.cases import problem, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _ops(p, kind):
    from trot_amd import meas, trial
    from trot_amd.core import HamChol, System

    N, na = p["h1"].shape[0], p["ca"].shape[1]
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    if kind == "uhf":
        nb = p["cb"].shape[1]
        sys_ = System(N, (na, nb), "unrestricted")
        tr = trial.UhfTrial(p["ca"], p["cb"])
        mops = meas.make_uhf_meas_ops(sys_)
        ref = O.make_uhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], p["cb"])
        w = (p["wa"], p["wb"])
    else:
        nb = p["cb"].shape[1] if "cb" in p else na
        sys_ = System(N, (na, nb), kind)
        tr = trial.RhfTrial(p["ca"])
        mops = meas.make_rhf_meas_ops(sys_)
        ref = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], kind)
        w = (p["wa"], p["wb"]) if kind == "unrestricted" else p["wa"]
    return sys_, ham, tr, mops, ref, w


def _dev(w):
    if isinstance(w, tuple):
        return tuple(torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in w)
    return torch.from_numpy(np.ascontiguousarray(w)).cuda()


@pytest.mark.parametrize("kind,N,n,W", [
    ("restricted", 21, 5, 7),
    ("restricted", 300, 7, 3),
    ("unrestricted", 26, 5, 5),
    ("uhf", 19, 4, 6),
])
def test_observable_kernels_match_oracle(gpu, kind, N, n, W):
    print(f"[observables] kernels {kind} N={N} n={n} W={W}", flush=True)
    p = problem(N, n, 8, W, seed=71, nocc_b=(n - 1 if kind == "uhf" else n if kind == "unrestricted" else None))
    sys_, ham, tr, mops, ref, w = _ops(p, kind)
    mctx = mops.build_meas_ctx(ham, tr)
    for name in ("rdm1", "density_corr"):
        got = mops.require_observable(name)(_dev(w), ham, mctx, tr).cpu().numpy()
        want = ref.observables[name](w)
        assert got.shape == want.shape, name
        assert rel(got, want) < 1e-12, name
    # one walker without the batch axis
    w0 = tuple(x[0] for x in w) if isinstance(w, tuple) else w[0]
    got = mops.require_observable("rdm1")(_dev(w0), ham, mctx, tr).cpu().numpy()
    assert rel(got, ref.observables["rdm1"](tuple(x[None] for x in w0) if isinstance(w0, tuple) else w0[None])[0]) \
        < 1e-12


@pytest.mark.parametrize("W,shape,cplx", [(37, (3, 5, 5), True), (1000, (20000,), False), (2, (1,), True)])
def test_weighted_mean(gpu, W, shape, cplx):
    from trot_amd import _lib

    print(f"[observables] weighted mean W={W} shape={shape}", flush=True)
    rng = np.random.default_rng(W)
    x = rng.standard_normal((W,) + shape)
    if cplx:
        x = x + 1j * rng.standard_normal((W,) + shape)
    wt = rng.uniform(0.0, 2.0, W)
    xd, wd = torch.from_numpy(x).cuda(), torch.from_numpy(wt).cuda()
    ws = torch.tensor([wt.sum()], dtype=torch.float64, device="cuda")
    a = _lib.weighted_mean(xd, wd, ws)
    b = _lib.weighted_mean(xd, wd, ws)
    assert torch.equal(a, b)
    assert rel(a.cpu().numpy(), O.weighted_block_mean(x, wt)) < 1e-13
    z = _lib.weighted_mean(xd, wd, torch.zeros(1, dtype=torch.float64, device="cuda"))
    assert not torch.count_nonzero(z).item()


@pytest.mark.parametrize("kind", ["restricted", "unrestricted"])
def test_block_observables_match_oracle(gpu, kind):
    from trot_amd import blocks, prop, trial

    print(f"[observables] block {kind}", flush=True)
    N, n, G, W = 22, 4, 26, 20
    p = problem(N, n, G, W, seed=72, nocc_b=(n if kind == "unrestricted" else None))
    sys_, ham, tr, mops, ref_ops, w = _ops(p, kind)
    tops = trial.make_rhf_trial_ops(sys_)
    mctx = mops.build_meas_ctx(ham, tr)
    pops = prop.make_prop_ops("restricted", kind)
    params = prop.QmcParams(dt=0.01, n_walkers=W, seed=3, n_prop_steps=3)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    ref_ctx = O.build_prop_ctx(p["h0"], p["h1"], p["chol"], O.rhf_rdm1(p["ca"]), 0.01)
    st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=_dev(w))
    ref = O.init_prop_state(w, ref_ops, W)
    fseq = p["fields"][:3]
    seed, ctr = st.rng_key
    zeta = float(pctx.fields.uniform((seed, ctr + 3))[0])
    names = ("rdm1", "density_corr")
    st, obs = blocks.block(st, sys=sys_, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                           meas_ctx=mctx, prop_ops=pops, prop_ctx=pctx, observable_names=names,
                           fields_seq=[torch.from_numpy(f).cuda() for f in fseq])
    ref, robs = O.block(ref, fseq, zeta, ref_ctx, ref_ops, O.QmcParams(dt=0.01, n_walkers=W, n_prop_steps=3),
                        observable_names=names)
    assert abs(obs.scalars["energy"].item() - robs["energy"]) < TOL * abs(robs["energy"])
    for name in names:
        assert rel(obs.observables[name].cpu().numpy(), robs["observables"][name]) < TOL, name
    dm = obs.observables["rdm1"].cpu().numpy()
    assert np.allclose(np.trace(dm, axis1=1, axis2=2), [n, n], atol=1e-10)
    with pytest.raises(KeyError):
        blocks.block(st, sys=sys_, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                     meas_ctx=mctx, prop_ops=pops, prop_ctx=pctx, observable_names=("spin_corr",))


def _graph_setup():
    from trot_amd import meas, prop, trial
    from trot_amd.core import HamChol, System

    p = problem(24, 6, 40, 16, seed=31)
    sys_ = System(24, (6, 6), "restricted")
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    mops = meas.make_rhf_meas_ops(sys_)
    tops = trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted")
    params = prop.QmcParams(dt=0.01, n_walkers=16, seed=11, n_prop_steps=3, n_blocks=30, n_eql_blocks=2)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    w0 = torch.from_numpy(p["wa"]).cuda()
    kw = dict(sys=sys_, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
              prop_ops=pops, prop_ctx=pctx)

    def init():
        return pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                    params=params, initial_walkers=w0)

    return kw, init


def test_graph_block_observables_match_eager(gpu):
    from trot_amd.blocks import block
    from trot_amd.graph import GraphBlock

    print("[observables] graph block", flush=True)
    kw, init = _graph_setup()
    names = ("rdm1", "density_corr")
    s_e = init()
    eager = []
    for _ in range(3):
        s_e, o = block(s_e, observable_names=names, **kw)
        eager.append(o)
    gb = GraphBlock()
    s_g = init()
    for b in range(3):
        s_g, o = gb(s_g, observable_names=names, **kw)
        torch.cuda.synchronize()
        assert torch.equal(o.scalars["energy"], eager[b].scalars["energy"]), b
        for name in names:
            assert torch.equal(o.observables[name], eager[b].observables[name]), (b, name)
        torch.randn(4096, device="cuda").sum().item()
    with pytest.raises(ValueError):
        gb(s_g, observable_names=("rdm1",), **kw)


def test_run_qmc_collects_observables(gpu):
    from trot_amd import blocks, driver, stats

    print("[observables] run_qmc", flush=True)
    kw, init = _graph_setup()
    params = kw["params"]
    res = driver.run_qmc(sys=kw["sys"], params=params, ham_data=kw["ham_data"], trial_data=kw["trial_data"],
                         meas_ops=kw["meas_ops"], trial_ops=kw["trial_ops"], prop_ops=kw["prop_ops"],
                         block_fn=blocks.block, state=init(), meas_ctx=kw["meas_ctx"], prop_ctx=kw["prop_ctx"],
                         observable_names=("rdm1",), verbose=False)
    obs = res.block_observables["rdm1"]
    n_kept = res.block_energies.shape[0] - 1 - params.n_eql_blocks
    assert obs.shape == (params.n_eql_blocks + n_kept, 2, 24, 24)
    smp, w = obs[params.n_eql_blocks:], res.block_weights[1 + params.n_eql_blocks:]
    mean = res.observable_means["rdm1"]
    assert np.array_equal(mean, driver.weighted_block_mean(smp, w))
    assert np.allclose(np.trace(mean, axis1=1, axis2=2).real, [6, 6], atol=1e-10)
    st = stats.blocking_analysis_ratio(res.block_energies[1 + params.n_eql_blocks:], w)
    num, den = stats.rebin_observable(smp, w, st["B_star"])
    err = res.observable_stderrs["rdm1"]
    assert err.shape == (2, 24, 24)
    if num.shape[0] >= 2:
        assert np.array_equal(err, stats.jackknife_ratios(num, den)[1]) and np.isfinite(err).all()
    else:
        assert np.isnan(err).all()


def test_run_blocks_and_run_qmc_energy(gpu):
    """make_run_blocks (trot/driver.py:64-110) and run_qmc_energy (trot/driver.py:771-803)."""
    from trot_amd import blocks, driver

    print("[observables] make_run_blocks / run_qmc_energy", flush=True)
    kw, init = _graph_setup()
    run_blocks = driver.make_run_blocks(block_fn=blocks.block, sys=kw["sys"], params=kw["params"],
                                        trial_ops=kw["trial_ops"], meas_ops=kw["meas_ops"], prop_ops=kw["prop_ops"],
                                        observable_names=("rdm1",))
    st, scal, obs = run_blocks(init(), ham_data=kw["ham_data"], trial_data=kw["trial_data"],
                               meas_ctx=kw["meas_ctx"], prop_ctx=kw["prop_ctx"], n_blocks=3)
    assert scal["energy"].shape == (3,) and scal["weight"].shape == (3,) and obs[0].shape == (3, 2, 24, 24)
    s_e = init()
    for b in range(3):
        s_e, o = blocks.block(s_e, observable_names=("rdm1",), **kw)
        assert torch.equal(o.scalars["energy"], scal["energy"][b]) and torch.equal(o.observables["rdm1"], obs[0][b])
    assert torch.equal(s_e.walkers, st.walkers)
    args = dict(sys=kw["sys"], params=kw["params"], ham_data=kw["ham_data"], trial_data=kw["trial_data"],
                meas_ops=kw["meas_ops"], trial_ops=kw["trial_ops"], prop_ops=kw["prop_ops"], block_fn=blocks.block,
                meas_ctx=kw["meas_ctx"], prop_ctx=kw["prop_ctx"], verbose=False)
    mean, err, be, bw = driver.run_qmc_energy(state=init(), **args)
    res = driver.run_qmc(state=init(), **args)
    assert mean == res.mean_energy and err == res.stderr_energy
    assert np.array_equal(be, res.block_energies) and np.array_equal(bw, res.block_weights)
