"""The allocation-free comb (trot_sr_comb_dev_noalloc) against trot_sr_comb_dev, through the C ABI.

The no-alloc entry point keeps its cumulative weights in a static device array that persists between
calls; two calls in a row, the second with fewer walkers, must each give the same bits as the
workspace form.  Sizes beyond the static limit are refused before any GPU work.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_W = 262144  # TROT_SR_NOALLOC_MAX_W in include/trot_amd.h


def _comb_pair(gpu, weights_np: np.ndarray, zeta: float):
    import torch

    from trot_amd import _lib

    h = _lib.lib()
    W = weights_np.shape[0]
    w = torch.from_numpy(weights_np).to(gpu)
    z = torch.tensor([zeta], dtype=torch.float64, device=gpu)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    idx_ref = torch.full((W,), -1, dtype=torch.int32, device=gpu)
    wn_ref = torch.full((W,), -1.0, dtype=torch.float64, device=gpu)
    ws = torch.empty(int(h.trot_sr_workspace(W)), dtype=torch.uint8, device=gpu)
    rc = h.trot_sr_comb_dev(ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(z.data_ptr()), W,
                            ctypes.c_void_p(idx_ref.data_ptr()), ctypes.c_void_p(wn_ref.data_ptr()),
                            ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream)
    assert rc == 0, h.trot_last_error()

    idx = torch.full((W,), -1, dtype=torch.int32, device=gpu)
    wn = torch.full((W,), -1.0, dtype=torch.float64, device=gpu)
    rc = h.trot_sr_comb_dev_noalloc(ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(z.data_ptr()), W,
                                    ctypes.c_void_p(idx.data_ptr()), ctypes.c_void_p(wn.data_ptr()), stream)
    assert rc == 0, h.trot_last_error()
    torch.cuda.synchronize()
    return idx.cpu().numpy(), wn.cpu().numpy(), idx_ref.cpu().numpy(), wn_ref.cpu().numpy()


def test_noalloc_comb_matches_workspace_comb_on_two_calls(gpu):
    rng = np.random.default_rng(11)
    # first call: many walkers, mixed signs and zeros (the comb uses |w|)
    w1 = rng.standard_normal(5000) * rng.uniform(0.0, 3.0, 5000)
    w1[::17] = 0.0
    idx, wn, idx_ref, wn_ref = _comb_pair(gpu, w1, 0.3141)
    assert np.array_equal(idx, idx_ref)
    assert np.array_equal(wn.view(np.int64), wn_ref.view(np.int64))
    assert idx.min() >= 0 and idx.max() < w1.size
    # second call, fewer walkers: stale cumulative weights of the first call sit past W in the scratch
    w2 = rng.uniform(0.0, 2.0, 777)
    idx, wn, idx_ref, wn_ref = _comb_pair(gpu, w2, 0.9)
    assert np.array_equal(idx, idx_ref)
    assert np.array_equal(wn.view(np.int64), wn_ref.view(np.int64))
    # the comb itself: w_new = sum|w| / W, and walker k is drawn about |w_k| W / sum|w| times
    assert np.allclose(wn, np.sum(np.abs(w2)) / w2.size, rtol=1e-14, atol=0)
    counts = np.bincount(idx, minlength=w2.size)
    expect = np.abs(w2) * w2.size / np.sum(np.abs(w2))
    assert np.all(np.abs(counts - expect) <= 1.0 + 1e-9)


def test_noalloc_comb_at_the_size_limit_and_beyond(gpu):
    from trot_amd import _lib

    h = _lib.lib()
    rng = np.random.default_rng(12)
    w = rng.uniform(0.1, 1.0, MAX_W)
    idx, wn, idx_ref, wn_ref = _comb_pair(gpu, w, 0.5)
    assert np.array_equal(idx, idx_ref)
    assert np.array_equal(wn.view(np.int64), wn_ref.view(np.int64))
    d = ctypes.c_void_p(16)  # never dereferenced: the size check returns first
    assert h.trot_sr_comb_dev_noalloc(d, d, MAX_W + 1, d, d, None) == -2
    assert b"TROT_SR_NOALLOC_MAX_W" in h.trot_last_error()
