"""CISD trial / measurement factories: the reference's configuration errors (no GPU needed)."""

from __future__ import annotations

import pytest

from trot_amd import cisd
from trot_amd.core import System


def test_trial_ops_errors():
    with pytest.raises(ValueError):
        cisd.make_cisd_trial_ops(System(10, (3, 2), "restricted"))  # trot/trial/cisd.py:190-191
    with pytest.raises(ValueError):
        cisd.make_cisd_trial_ops(System(10, (3, 3), "unrestricted"))  # :192-195
    with pytest.raises(ValueError):
        cisd.make_cisd_trial_ops(System(10, (3, 3), "restricted"), memory_mode="medium")  # :196-197
    ops = cisd.make_cisd_trial_ops(System(10, (3, 3), "restricted"), memory_mode="low")
    assert ops.overlap is cisd.overlap_r and ops.get_rdm1 is cisd.get_rdm1


def test_meas_ops_config():
    with pytest.raises(ValueError):
        cisd.make_cisd_meas_ops(System(10, (3, 3), "unrestricted"))  # trot/meas/cisd.py:829-832
    ops = cisd.make_cisd_meas_ops(System(10, (3, 3), "restricted"))
    cfg = cisd.get_cisd_meas_cfg(ops)
    assert cfg.memory_mode == "high" and cfg.force_bias_precision == "bf16x3"
    assert ops.engine.precision == "bf16x3" and ops.has_kernel("force_bias") and ops.has_kernel("energy")
    assert ops.has_observable("rdm1")
    ops64 = cisd.make_cisd_meas_ops(System(10, (3, 3), "restricted"), mixed_precision=False, testing=True)
    assert cisd.get_cisd_meas_cfg(ops64).force_bias_precision == "f64" and ops64.engine.precision == "f64"
    assert ops.engine.doubles_precision == "bf16x3" and ops64.engine.doubles_precision == "f64"
    from trot_amd import meas

    assert cisd.get_cisd_meas_cfg(meas.make_rhf_meas_ops(System(10, (3, 3), "restricted"))) is None
