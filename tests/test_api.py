"""Host API mirrors the reference's factories, defaults and error behaviour (CPU only)."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from trot_amd import meas, prop, trial
from trot_amd.core import HamChol, MeasOps, System


def test_qmc_params_defaults_match_reference():
    """trot/prop/types.py:30-47."""
    p = prop.QmcParams(seed=1)
    assert (p.dt, p.n_chunks, p.n_exp_terms, p.n_prop_steps, p.n_blocks, p.n_walkers) == (0.005, 1, 6, 50, 200, 200)
    assert (p.pop_control_damping, p.weight_floor, p.weight_cap, p.shift_ema, p.n_eql_blocks) == (
        0.1, 1e-3, 100.0, 0.1, 20)
    with pytest.raises(dataclasses.FrozenInstanceError):
        p.dt = 0.1


def test_make_prop_ops_errors():
    """trot/prop/chol_afqmc_ops.py:259-287."""
    with pytest.raises(ValueError):
        prop.make_prop_ops("bogus", "restricted")
    with pytest.raises(ValueError):
        prop.make_prop_ops("restricted", "bogus")
    with pytest.raises(NotImplementedError):
        prop.make_prop_ops("restricted", "generalized")
    with pytest.raises(ValueError):
        prop.make_prop_ops("restricted", "restricted", precision="f16")
    ops = prop.make_prop_ops("restricted", "Restricted", mixed_precision=True)
    assert callable(ops.step) and callable(ops.build_prop_ctx) and callable(ops.init_prop_state)


def test_trial_and_meas_factories():
    """trot/trial/rhf.py:66-87, trot/meas/rhf.py:322-367, trot/core/ops.py:148-160."""
    with pytest.raises(ValueError):
        trial.make_rhf_trial_ops(System(4, (2, 1), "restricted"))
    with pytest.raises(ValueError):
        meas.make_rhf_meas_ops(System(4, (2, 2), "restricted"), memory_mode="medium")
    with pytest.raises(ValueError):
        meas.make_rhf_meas_ops(System(4, (2, 2), "weird"))
    m = meas.make_rhf_meas_ops(System(4, (2, 2), "unrestricted"), memory_mode="low")
    assert m.has_kernel("force_bias") and m.has_kernel("energy")
    assert m.available_kernels() == ("energy", "force_bias")
    with pytest.raises(KeyError, match="missing required kernel 'nope'"):
        m.require_kernel("nope")
    with pytest.raises(NotImplementedError):
        meas.make_uhf_meas_ops(System(4, (2, 1), "restricted"))
    assert isinstance(meas.make_uhf_meas_ops(System(4, (2, 1), "unrestricted")), MeasOps)


def test_system_and_ham_containers():
    s = System(6, (3, 2), "UnRestricted")
    assert s.walker_kind == "unrestricted" and (s.nup, s.ndn, s.ne) == (3, 2, 5)
    h = HamChol(0.0, np.zeros((3, 3)), np.zeros((4, 3, 3)))
    assert h.nchol == 4
    with pytest.raises(ValueError):
        HamChol(0.0, np.zeros((3, 3)), np.zeros((4, 3, 3)), nchol=5)
    with pytest.raises(ValueError):
        HamChol(0.0, np.zeros((3, 3)), np.zeros((4, 3, 3)), basis="spin")


def test_complex_trial_rejected():
    import torch

    if not torch.cuda.is_available():
        with pytest.raises((ValueError, RuntimeError, AssertionError)):
            trial.RhfTrial(np.eye(3)[:, :1] * (1 + 1j))
        return
    with pytest.raises(ValueError):
        trial.RhfTrial(np.eye(3)[:, :1] * (1 + 1j))


def test_driver_and_block_names_of_the_reference():
    """The callables a trot driver script imports from trot.driver, trot.prop.blocks and trot.stat_utils
    exist under the same names (trot/driver.py:64-110, 168-406, 771-803; trot/prop/blocks.py:66-195;
    trot/stat_utils.py:45-313)."""
    import inspect

    from trot_amd import blocks, driver, stats

    for mod, names in ((driver, ["make_run_blocks", "run_qmc", "run_qmc_energy", "QmcResult"]),
                       (blocks, ["block", "BlockObs", "dump_prop_state_npz", "load_prop_state_npz",
                                 "load_all_prop_states_npz", "make_block_state_logger"]),
                       (stats, ["blocking_analysis_ratio", "reject_outliers", "jackknife_ratios",
                                "rebin_observable"])):
        for n in names:
            assert hasattr(mod, n), f"{mod.__name__}.{n}"
    assert list(driver.QmcResult._fields) == ["mean_energy", "stderr_energy", "block_energies", "block_weights",
                                              "block_observables", "observable_means", "observable_stderrs"]
    sig = inspect.signature(driver.run_qmc).parameters
    for p in ("sys", "params", "ham_data", "trial_data", "meas_ops", "trial_ops", "prop_ops", "block_fn", "state",
              "meas_ctx", "prop_ctx", "target_error", "mesh", "observable_names"):
        assert p in sig, p
