"""The device field stream and the graph-recorded block driver.

* Philox4x64-10 on the GPU: raw words bit for bit against the CPU twin (oracle/philox.py, itself
  pinned to numpy.random.Philox), normals against the twin's Box-Muller to 1e-13 (libm rounding
  of log / sincos), the comb offset bit for bit, device-counter mode equal to by-value mode;
* SR with the comb offset read on the device equals SR with it passed by value;
* GraphStep (one HIP graph per step) against eager steps, bit for bit, FP64 and mixed;
* GraphBlock (one HIP graph per block) against the eager block with the same device stream: every
  state tensor and both block scalars bit-identical over three blocks (a warm-up block and two
  replays, with unrelated eager GPU work and host synchronisation between replays), for restricted,
  unrestricted-UHF and the mixed split-BF16 path; and run_qmc with it against run_qmc eager.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import philox as P
from tests.cases import problem

pytestmark = pytest.mark.gpu


def _u64(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("seed,counter,purpose,nb", [(0, 0, 0, 1), (12345, 7, 0, 1000), (2**64 - 1, 2**63 + 1, 1, 37)])
def test_philox_raw_bit_exact(gpu, seed, counter, purpose, nb):
    from trot_amd import _lib

    print(f"[philox raw] seed {seed} counter {counter} purpose {purpose} blocks {nb}", flush=True)
    g = _u64(_lib.philox_raw(seed, counter, purpose, nb))
    assert np.array_equal(g, P.raw_blocks(seed, counter, purpose, nb))


@pytest.mark.parametrize("n", [1, 6, 4 * 513, 100_003])
def test_philox_normal_and_uniform(gpu, n):
    from trot_amd import _lib

    seed, ctr = 77, 5
    z = _lib.philox_normal((n,), seed=seed, counter=ctr)
    ref = P.normals(seed, ctr, n)
    assert np.allclose(z.cpu().numpy(), ref, rtol=1e-13, atol=1e-13)
    st = _lib.rng_state(seed, ctr)
    z2 = _lib.philox_normal((n,), state=st)
    assert torch.equal(z, z2)
    _lib.rng_advance(st, 3)
    assert _u64(st).tolist() == [seed, ctr + 3]
    assert torch.equal(_lib.philox_normal((n,), state=st), _lib.philox_normal((n,), seed=seed, counter=ctr + 3))
    u = _lib.philox_uniform(state=st)
    assert u.item() == P.uniform(seed, ctr + 3)
    assert _lib.philox_uniform(seed=seed, counter=ctr + 3).item() == u.item()


def test_sr_with_device_offset(gpu):
    from trot_amd import _lib

    w = torch.from_numpy(np.random.default_rng(3).uniform(0.1, 2.0, 333)).cuda()
    zeta = _lib.philox_uniform(seed=9, counter=2)
    i1, w1 = _lib.sr_comb(w, zeta)
    i2, w2 = _lib.sr_comb(w, float(zeta.item()))
    assert torch.equal(i1, i2) and torch.equal(w1, w2)


def _setup(kind, precision, N, n, G, W, nps=3):
    from trot_amd import meas, prop, trial
    from trot_amd.core import HamChol, System

    uhf = kind == "uhf"
    p = problem(N, n, G, W, seed=31, nocc_b=(n - 1 if uhf else None))
    nb = n - 1 if uhf else n
    sys = System(N, (n, nb), "unrestricted" if uhf else kind)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    if uhf:
        tr = trial.UhfTrial(p["ca"], p["cb"])
        mops = meas.make_uhf_meas_ops(sys)
        tops = trial.make_uhf_trial_ops(sys)
    else:
        tr = trial.RhfTrial(p["ca"])
        mops = meas.make_rhf_meas_ops(sys, precision=precision)
        tops = trial.make_rhf_trial_ops(sys)
    pops = prop.make_prop_ops("restricted", sys.walker_kind, precision=precision,
                              taylor_precision=(precision if precision == "bf16x3" else None))
    params = prop.QmcParams(dt=0.01, n_walkers=W, seed=11, n_prop_steps=nps, n_blocks=4, n_eql_blocks=2)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    if uhf:
        w0 = (torch.from_numpy(p["wa"]).cuda(), torch.from_numpy(p["wb"]).cuda())
    else:
        w0 = torch.from_numpy(p["wa"]).cuda()
    kw = dict(sys=sys, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
              prop_ops=pops, prop_ctx=pctx)

    def init():
        return pops.init_prop_state(sys=sys, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                    params=params, initial_walkers=w0)

    return kw, init


def _flat(state):
    from trot_amd.batch import _tensors

    return _tensors(state)


@pytest.mark.parametrize("kind,precision,shape", [
    ("restricted", "f64", (24, 6, 40, 16)),
    ("uhf", "f64", (26, 5, 30, 12)),
    ("restricted", "bf16x3", (40, 6, 50, 24)),
])
def test_graph_block_matches_eager(gpu, kind, precision, shape):
    from trot_amd.blocks import block
    from trot_amd.graph import GraphBlock

    print(f"[graph block] {kind} {precision} N,n,G,W={shape}", flush=True)
    kw, init = _setup(kind, precision, *shape)
    s_e = init()
    eager = []
    for _ in range(3):
        s_e, o = block(s_e, **kw)
        eager.append((s_e, o))
    gb = GraphBlock()
    s_g = init()
    for b in range(3):
        s_g, o = gb(s_g, **kw)
        torch.cuda.synchronize()
        se, oe = eager[b]
        assert s_g.rng_key == se.rng_key
        for x, y in zip(_flat(s_g), _flat(se)):
            assert torch.equal(x, y), b
        assert torch.equal(o.scalars["energy"], oe.scalars["energy"]), b
        assert torch.equal(o.scalars["weight"], oe.scalars["weight"]), b
        # unrelated eager work and a host round trip between replays
        torch.randn(4096, device="cuda").sum().item()
    assert gb.blocks_done == 3
    # a state from outside is copied in: replaying from the eager run's second block gives its third
    s_x, o_x = gb(eager[1][0], **kw)
    for x, y in zip(_flat(s_x), _flat(eager[2][0])):
        assert torch.equal(x, y)
    assert torch.equal(o_x.scalars["energy"], eager[2][1].scalars["energy"])


@pytest.mark.parametrize("precision", ["f64", "bf16x3"])
def test_graph_step_matches_eager(gpu, precision):
    from trot_amd.graph import GraphStep

    print(f"[graph step] {precision}", flush=True)
    kw, init = _setup("restricted", precision, 40, 6, 50, 24)
    step_kw = {k: v for k, v in kw.items() if k not in ("sys", "prop_ops")}
    pops = kw["prop_ops"]
    s_e = init()
    eager = []
    for _ in range(4):
        s_e = pops.step(s_e, **step_kw)
        eager.append(s_e)
    gs = GraphStep(pops)
    s_g = init()
    for i in range(4):
        s_g = gs(s_g, **step_kw)
        torch.cuda.synchronize()
        assert s_g.rng_key == eager[i].rng_key
        for x, y in zip(_flat(s_g), _flat(eager[i])):
            assert torch.equal(x, y), i
        torch.randn(4096, device="cuda").sum().item()


def test_run_qmc_graph_equals_eager(gpu):
    from trot_amd import driver
    from trot_amd.graph import GraphBlock

    kw, init = _setup("restricted", "f64", 24, 6, 40, 16, nps=2)
    args = dict(sys=kw["sys"], params=kw["params"], ham_data=kw["ham_data"], trial_data=kw["trial_data"],
                meas_ops=kw["meas_ops"], trial_ops=kw["trial_ops"], prop_ops=kw["prop_ops"],
                meas_ctx=kw["meas_ctx"], prop_ctx=kw["prop_ctx"], verbose=False)
    r_e = driver.run_qmc(state=init(), **args)
    r_g = driver.run_qmc(state=init(), block_fn=GraphBlock(), **args)
    assert np.array_equal(r_e.block_energies, r_g.block_energies)
    assert np.array_equal(r_e.block_weights, r_g.block_weights)
    assert r_e.mean_energy == r_g.mean_energy


def test_graph_block_free_projection(gpu):
    """GraphBlock(block_fp) and run_qmc_fp with it against the eager free-projection block and driver."""
    from trot_amd import fp, meas, trial
    from trot_amd.core import HamChol, System
    from trot_amd.graph import GraphBlock

    N, n, G, W = 24, 5, 30, 12
    p = problem(N, n, G, W, seed=61)
    sys = System(N, (n, n), "restricted")
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    mops = meas.make_rhf_meas_ops(sys)
    tops = trial.make_rhf_trial_ops(sys)
    params = fp.QmcParamsFp(dt=0.01, n_walkers=W, seed=5, n_prop_steps=4, n_qr_blocks=2, n_blocks=3, n_traj=2,
                            ene0=-1.0)
    pops = fp.make_prop_ops_fp("restricted", "restricted", sys)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    kw = dict(sys=sys, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
              prop_ops=pops, prop_ctx=pctx)
    w0 = torch.from_numpy(p["wa"]).cuda()

    def init():
        return pops.init_prop_state(sys=sys, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                    params=params, initial_walkers=w0)

    s_e = init()
    eager = []
    for _ in range(3):
        s_e, o = fp.block_fp(s_e, **kw)
        eager.append((s_e, o))
    gb = GraphBlock(fp.block_fp, fp.draws_per_block)
    s_g = init()
    for b in range(3):
        s_g, o = gb(s_g, **kw)
        torch.cuda.synchronize()
        se, oe = eager[b]
        assert s_g.rng_key == se.rng_key
        for x, y in zip(_flat(s_g), _flat(se)):
            assert torch.equal(x, y), b
        for k in ("energy", "weight", "overlap", "abs_overlap"):
            assert torch.equal(o.scalars[k], oe.scalars[k]), (b, k)
        torch.randn(4096, device="cuda").sum().item()
    args = dict(sys=sys, params=params, ham_data=ham, trial_data=tr, meas_ops=mops, trial_ops=tops, prop_ops=pops,
                meas_ctx=mctx, prop_ctx=pctx, verbose=False)
    r_e = fp.run_qmc_fp(state=init(), **args)
    r_g = fp.run_qmc_fp(state=init(), block_fn=GraphBlock(fp.block_fp, fp.draws_per_block), **args)
    assert np.array_equal(r_e.block_energies, r_g.block_energies)
    assert np.array_equal(r_e.mean_energy, r_g.mean_energy)
