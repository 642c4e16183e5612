"""Per-block state files (trot/prop/blocks.py:66-192; the reference's tests/test_block_state_logger.py).

CPU: ``dump_prop_state_npz`` / ``load_prop_state_npz`` round-trip every field bit for bit, for
restricted and unrestricted walkers.  GPU: ``run_qmc`` with ``make_block_state_logger`` writes one file per
equilibration and sampling block, and a block run from a loaded file (its Green's functions recomputed
from the walkers) reproduces the next file bit for bit: a run can resume from any block.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.cases import problem
from trot_amd import blocks
from trot_amd.prop import PropState


def _state(rng, W, N, spins):
    def c(*shape):
        return torch.from_numpy(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))

    walkers = tuple(c(W, N, n) for n in spins) if len(spins) > 1 else c(W, N, spins[0])
    return PropState(walkers=walkers, weights=torch.from_numpy(rng.uniform(0, 2, W)), overlaps=c(W),
                     rng_key=(7, 123456789012), pop_control_ene_shift=torch.tensor(-1.25, dtype=torch.float64),
                     e_estimate=torch.tensor(-1.5, dtype=torch.float64),
                     node_encounters=torch.tensor(3, dtype=torch.int64))


@pytest.mark.parametrize("spins", [(3,), (3, 2)])
@pytest.mark.parametrize("compressed", [True, False])
def test_state_file_round_trip(tmp_path, spins, compressed):
    st = _state(np.random.default_rng(len(spins)), 5, 6, spins)
    blocks.dump_prop_state_npz(st, tmp_path / "s.npz", compressed=compressed)
    back = blocks.load_prop_state_npz(tmp_path / "s.npz", device="cpu")
    assert back.rng_key == st.rng_key and back.green is None
    a = st.walkers if isinstance(st.walkers, tuple) else (st.walkers,)
    b = back.walkers if isinstance(back.walkers, tuple) else (back.walkers,)
    assert isinstance(back.walkers, tuple) == isinstance(st.walkers, tuple)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    for name in ("weights", "overlaps", "pop_control_ene_shift", "e_estimate", "node_encounters"):
        assert torch.equal(getattr(st, name), getattr(back, name)), name
    with np.load(tmp_path / "s.npz") as d:
        assert sorted(d.files) == sorted(["weights", "overlaps", "rng_key", "pop_control_ene_shift", "e_estimate",
                                          "node_encounters"] + (["walkers"] if len(spins) == 1 else
                                                                ["walkers_0", "walkers_1"]))


def test_load_all_in_file_name_order(tmp_path):
    rng = np.random.default_rng(0)
    states = [_state(rng, 2, 3, (1,)) for _ in range(3)]
    for i, st in enumerate(states):
        blocks.dump_prop_state_npz(st, tmp_path / f"block_state_{i:05d}.npz")
    blocks.dump_prop_state_npz(states[0], tmp_path / "other_00000.npz")
    got = blocks.load_all_prop_states_npz(tmp_path, device="cpu")
    assert len(got) == 3
    for st, g in zip(states, got):
        assert torch.equal(st.walkers, g.walkers)


@pytest.mark.gpu
def test_logger_writes_each_block_and_a_run_resumes_from_any(gpu, tmp_path):
    from trot_amd import driver, meas, prop, trial
    from trot_amd.core import HamChol, System

    print("[block states] logger and resume", flush=True)
    p = problem(10, 3, 12, 8, seed=81)
    sys_ = System(10, (3, 3), "restricted")
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    mops = meas.make_rhf_meas_ops(sys_)
    tops = trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted")
    params = prop.QmcParams(dt=0.01, n_walkers=8, seed=11, n_prop_steps=2, n_eql_blocks=2, n_blocks=3)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    st0 = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                               initial_walkers=torch.from_numpy(p["wa"]).cuda())
    res = driver.run_qmc(sys=sys_, params=params, ham_data=ham, trial_data=tr, meas_ops=mops, trial_ops=tops,
                         prop_ops=pops, block_fn=blocks.make_block_state_logger(tmp_path), state=st0,
                         meas_ctx=mctx, prop_ctx=pctx, verbose=False)
    assert np.isfinite(res.mean_energy)
    files = sorted(tmp_path.glob("block_state_*.npz"))
    assert len(files) == params.n_eql_blocks + params.n_blocks
    states = blocks.load_all_prop_states_npz(tmp_path)
    assert states[0].walkers.shape == (8, 10, 3) and states[-1].weights.shape == (8,)
    kw = dict(sys=sys_, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
              prop_ops=pops, prop_ctx=pctx)
    for k in range(len(states) - 1):
        nxt, _ = blocks.block(states[k], **kw)
        ref = states[k + 1]
        assert nxt.rng_key == ref.rng_key, k
        for name in ("walkers", "weights", "overlaps", "pop_control_ene_shift", "e_estimate", "node_encounters"):
            assert torch.equal(getattr(nxt, name), getattr(ref, name)), (k, name)
