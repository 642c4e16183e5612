"""The staged-input HDF5 reader (trot_amd/staged.py) against files written by the HDF5 library itself.

tests/golden/make_staged_h5.py wrote the fixtures with h5py 3.3 / HDF5 1.10.6 in the reference's cache
layout (trot/staging.py:1353-1376), in both file-format generations h5py produces (superblock 0 with
symbol-table groups; superblock 3 with version-2 object headers and link messages), and saved the
arrays it wrote to staged_expected.npz.  Every array must come back bit for bit, every attribute with
its value and type, and the unsupported chunked / compressed layout must be refused, not misread.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"


def _exp():
    return np.load(GOLD / "staged_expected.npz")


def test_rhf_cache_superblock0():
    from trot_amd import staged

    exp = _exp()
    s = staged.load_staged(GOLD / "staged_rhf_earliest.h5")
    assert s.ham.h0 == float(exp["rhf_h0"])
    assert np.array_equal(s.ham.h1, exp["rhf_h1"]) and s.ham.h1.dtype == np.float64
    assert np.array_equal(s.ham.chol, exp["rhf_chol"]) and s.ham.chol.shape == (7, 6, 6)
    assert s.ham.nelec == (3, 3) and s.ham.norb == 6 and s.ham.chol_cut == 1e-5
    assert s.ham.frozen == 0 and isinstance(s.ham.frozen, int)
    assert s.ham.source_kind == "mf" and s.ham.basis == "restricted"
    assert s.trial.kind == "rhf" and s.trial.source_kind == "mf" and s.trial.frozen == 0
    assert list(s.trial.data) == ["mo"] and np.array_equal(s.trial.data["mo"], exp["rhf_mo"])
    assert s.meta == {"frozen": 0, "basis": "sto-6g", "note": "rhf fixture"}


def test_uhf_cache_superblock3_with_frozen_arrays():
    from trot_amd import staged

    exp = _exp()
    s = staged.load_staged(GOLD / "staged_uhf_latest.h5")
    assert s.ham.h0 == float(exp["uhf_h0"])
    assert np.array_equal(s.ham.chol, exp["uhf_chol"]) and np.array_equal(s.ham.h1, exp["uhf_h1"])
    assert s.ham.nelec == (3, 2) and s.ham.norb == 5
    assert np.array_equal(s.ham.frozen, exp["uhf_frozen"]) and s.ham.frozen.dtype == np.int64
    assert np.array_equal(s.trial.frozen, exp["uhf_frozen"])
    assert sorted(s.trial.data) == ["mo_a", "mo_b"]
    assert np.array_equal(s.trial.data["mo_a"], exp["uhf_mo_a"])
    assert np.array_equal(s.trial.data["mo_b"], exp["uhf_mo_b"])
    assert np.array_equal(s.meta["frozen"], [0, 7]) and s.meta["note"] == "uhf fixture"


def test_generic_access_and_refusals(tmp_path):
    from trot_amd import staged

    f = staged.H5File(GOLD / "staged_rhf_earliest.h5")
    assert sorted(f.keys("/")) == ["ham", "trial"]
    assert "ham/chol" in f and "ham/nothing" not in f
    with pytest.raises(TypeError):
        f["ham"]  # a group, not a dataset
    with pytest.raises(KeyError):
        f["trial/data/mo_a"]
    with pytest.raises(NotImplementedError):
        staged.load_staged(GOLD / "staged_chunked.h5")  # chunked + gzip: refused, not misread
    bad = tmp_path / "x.h5"
    bad.write_bytes(b"not hdf5" * 100)
    with pytest.raises(ValueError):
        staged.H5File(bad)


def test_staged_hamiltonian_object():
    from trot_amd import staged

    s = staged.load_staged(GOLD / "staged_uhf_latest.h5")
    ham = staged.ham_chol(s)
    assert ham.nchol == 9 and ham.basis == "restricted"
    assert np.array_equal(np.asarray(ham.chol), s.ham.chol)


@pytest.mark.gpu
def test_staged_rhf_cache_drives_the_gpu_step(gpu):
    """The cached Hamiltonian and trial go through make_rhf_trial_data / the FP64 step and match the oracle
    run on the arrays the HDF5 library wrote (two steps, identical fields)."""
    import torch

    from oracle import afqmc_oracle as O
    from tests.cases import rel
    from trot_amd import meas, prop, staged, trial
    from trot_amd.core import System

    s = staged.load_staged(GOLD / "staged_rhf_earliest.h5")
    sys_ = System(s.ham.norb, s.ham.nelec, "restricted")
    ham = staged.ham_chol(s).to_device()
    tr = trial.make_rhf_trial_data(s.trial.data, sys_)  # trot/trial/rhf.py:90-93
    mops = meas.make_rhf_meas_ops(sys_)
    tops = trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted")
    W = 8
    params = prop.QmcParams(dt=0.01, n_walkers=W, seed=2)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    c = s.trial.data["mo"][:, :3]
    rng = np.random.default_rng(5)
    w0 = c[None] + 0.05 * (rng.standard_normal((W, 6, 3)) + 1j * rng.standard_normal((W, 6, 3)))
    st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=torch.from_numpy(w0).cuda())
    ref_ops = O.make_rhf_ops(s.ham.h0, s.ham.h1, s.ham.chol, c)
    ref_ctx = O.build_prop_ctx(s.ham.h0, s.ham.h1, s.ham.chol, O.rhf_rdm1(c), 0.01)
    ref = O.init_prop_state(w0, ref_ops, W)
    ref_params = O.QmcParams(dt=0.01, n_walkers=W)
    for k in range(2):
        f = rng.standard_normal((W, s.ham.chol.shape[0]))
        st = pops.step(st, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
                       prop_ctx=pctx, fields=torch.from_numpy(f).cuda())
        ref = O.afqmc_step(ref, f, ref_ctx, ref_ops, ref_params)
    assert rel(st.walkers.cpu().numpy(), ref.walkers) < 1e-10
    assert rel(st.weights.cpu().numpy(), ref.weights) < 1e-10


def test_truncated_file_is_refused(tmp_path):
    """Contiguous storage that runs past the end of the file is an error (as in h5py), not a zero tail."""
    from trot_amd import staged

    data = (GOLD / "staged_rhf_earliest.h5").read_bytes()
    exp = _exp()
    start = data.find(exp["rhf_chol"].tobytes())
    assert start > 0
    cut = tmp_path / "truncated.h5"
    cut.write_bytes(data[:start + exp["rhf_chol"].nbytes // 2])  # ends inside the Cholesky vectors
    f = staged.H5File(cut)
    assert np.array_equal(f["ham/h1"], exp["rhf_h1"])  # datasets before the cut still read
    with pytest.raises(ValueError, match="truncated"):
        f["ham/chol"]
    with pytest.raises(ValueError, match="truncated"):
        staged.load_staged(cut)
