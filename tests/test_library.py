"""The C-ABI library: loads, exports every declared symbol, rejects bad input (no GPU needed).

The error paths return before any HIP call, so they run on a CPU-only host.
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "trot_amd.h"


def _declared():
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int|long|size_t)\s+(trot_\w+)\(", txt, re.M)))


def test_header_declares_the_entry_points():
    names = _declared()
    for must in ("trot_force_bias", "trot_vhs_build", "trot_taylor_apply", "trot_one_body_apply",
                 "trot_overlap_green", "trot_phaseless_update", "trot_local_energy", "trot_orthonormalize",
                 "trot_sr_comb", "trot_apply_trotter", "trot_field_shift", "trot_block_energy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from trot_amd import _lib

    h = _lib.lib()
    missing = [n for n in _declared() if not hasattr(h, n)]
    assert not missing, missing
    assert sorted(_lib.exported_symbols()) == _declared()
    assert b"gfx950" in h.trot_version()


def test_error_codes_without_gpu():
    from trot_amd import _lib

    h = _lib.lib()
    # null pointers
    assert h.trot_overlap_green(None, None, 4, 8, 2, None, None, None, 0, None) == -1
    assert b"null" in h.trot_last_error()
    # nocc above the kernels' limit
    dummy = ctypes.c_void_p(16)
    assert h.trot_overlap_green(dummy, dummy, 4, 300, 257, dummy, None, dummy, 1 << 30, None) == -2  # n > 256
    # workspace smaller than the query
    need = h.trot_overlap_green_workspace(4, 8, 2)
    assert need > 0
    assert h.trot_overlap_green(dummy, dummy, 4, 8, 2, dummy, None, dummy, need - 1, None) == -4
    assert b"workspace" in h.trot_last_error()
    assert h.trot_local_energy_workspace(8, 100) >= 8 * 4 * 16
    assert h.trot_local_energy(dummy, None, dummy, None, dummy, None, 4, 300, 257, 0, 3, 0.0, dummy, dummy, 1 << 30,
                               None) == -2
    assert h.trot_inverse_det(dummy, 4, 257, ctypes.c_void_p(32), dummy, None) == -2
    assert h.trot_force_bias(None, None, 1, 1, 1, 2.0, 0, None, None) == -1
    assert h.trot_sr_comb(dummy, 0.5, 0, dummy, dummy, dummy, 4096, None) == -1
    assert h.trot_sr_comb_dev_noalloc(dummy, None, 8, dummy, dummy, None) == -1
    assert h.trot_sr_comb_dev_noalloc(dummy, dummy, 262145, dummy, dummy, None) == -2  # above the static scratch
    assert h.trot_one_body_apply(dummy, dummy, 1, 4, 2, dummy, None) == -1  # in place is rejected
    assert h.trot_orthonormalize(dummy, 2, 8, 70, dummy, 1 << 30, None) == -1  # n > N
    assert h.trot_orthonormalize(dummy, 2, 300, 257, dummy, 1 << 30, None) == -2  # n > 256


def test_error_codes_of_mixed_and_batching_entry_points_without_gpu():
    """Argument checks of the split-BF16 propagation / energy, split-K force bias and tile tuner entry
    points return their codes before any GPU work is enqueued."""
    from trot_amd import _lib

    h = _lib.lib()
    d = ctypes.c_void_p(16)
    big = 1 << 40
    # trotter / Taylor on split planes
    assert h.trot_trotter_b3_symblk(None, d, d, d, 4, 8, 2, 0.1, 6, d, big, None) == -1
    assert h.trot_trotter_b3_symblk(d, d, d, d, 0, 8, 2, 0.1, 6, d, big, None) == -1
    assert h.trot_trotter_b3_symblk(d, d, d, d, 70000, 8, 2, 0.1, 6, d, big, None) == -1
    need = h.trot_taylor_b3_workspace(4, 8, 2)
    assert h.trot_trotter_b3_symblk(d, d, d, d, 4, 8, 2, 0.1, 6, d, need - 1, None) == -4
    assert h.trot_taylor_apply_b3_symblk(d, d, 4, 8, 2, 0.1, -1, d, big, None) == -1
    assert h.trot_onebody_b3_planes(None, 8, d, None) == -1
    assert h.trot_onebody_b3_planes_bytes(129) == 2 * 256 * 256 * 2  # rows and columns padded to 128
    # split-BF16 local energy
    assert h.trot_local_energy_b3(None, None, d, None, d, None, 4, 8, 2, 0, 3, 0.0, d, d, big, None) == -1
    assert h.trot_local_energy_b3(d, None, d, None, d, None, 4, 100, 65, 0, 3, 0.0, d, d, big, None) == -2
    need = h.trot_local_energy_b3_workspace(4, 8, 3, 0)
    assert need > 0
    assert h.trot_local_energy_b3(d, None, d, None, d, None, 4, 8, 2, 0, 3, 0.0, d, d, need - 1, None) == -4
    assert h.trot_rot_chol_b3(d, 3, 65, 8, d, None) == -2
    assert h.trot_rot_chol_b3_bytes(5, 2, 8) == 2 * (8 * 64) * 128 * 2  # G padded to 4 blocks of 64 rows
    # split-K force bias: slices for about 2048 workgroups (two of the 1280 tiles at W = 2048, G = 2560: partial
    # products of 2 W G complex numbers), none once the tiles alone reach that, slices for a small walker set
    assert h.trot_force_bias_workspace(2048, 2560, 512 * 64) == 2 * 2048 * 2560 * 16
    assert h.trot_force_bias_workspace(4096, 4096, 512 * 64) == 0
    assert h.trot_force_bias_workspace(128, 800, 200 * 40) > 0
    assert h.trot_force_bias_splitk(None, d, 4, 4, 4, 2.0, 0, d, None, 0, None) == -1
    assert h.trot_force_bias_splitk(d, d, 128, 800, 8000, 2.0, 0, d, None, 0, None) == -4
    # Taylor tile tuner and table
    assert h.trot_set_taylor_tile(4, 8, 2, 1, 7) == -1
    assert h.trot_autotune_taylor(0, 8, 2, 1, d, big, None, None, None) == -1
    assert h.trot_autotune_taylor(4, 8, 2, 1, d, 0, None, None, None) == -4
    assert h.trot_set_taylor_tile(3, 5, 2, 0, 3) == 0 and h.trot_get_taylor_tile(3, 5, 2, 0) == 3
    assert h.trot_get_taylor_tile(3, 5, 2, 1) == 0  # other storage: default 64 x 64
    # counter-based random numbers and SR with a device comb offset
    assert h.trot_philox_raw(0, 0, 0, 4, None, None) == -1
    assert h.trot_philox_raw(0, 0, 0, -1, d, None) == -1
    assert h.trot_philox_normal(None, 0, 0, 8, None, None) == -1
    assert h.trot_philox_normal(None, 0, 0, -2, d, None) == -1
    assert h.trot_philox_uniform(None, 0, 0, None, None) == -1
    assert h.trot_rng_advance(None, 1, None) == -1
    assert h.trot_sr_comb_dev(d, None, 4, d, d, d, big, None) == -1
    assert h.trot_sr_comb_dev(d, d, 4, d, d, d, 0, None) == -4
    # free projection
    assert h.trot_fp_prepare(None, d, 4, 8, 0.1, 0.01, 0.0, 0.0, 4, None, None, d, None) == -1
    assert h.trot_fp_prepare(d, d, 4, 8, 0.1, 0.01, 0.0, 0.0, 0, None, None, d, None) == -1
    assert h.trot_scale_rows(None, d, 4, 8, None) == -1
    assert h.trot_scale_rows(d, d, 0, 8, None) == 0  # nothing to scale
    assert h.trot_fp_weights(d, None, None, 4, d, d, None) == -1
    assert h.trot_fp_unnorm(d, d, None, 4, None) == -1
    assert h.trot_block_energy_fp(d, d, d, 4, 0.0, 0.0, d, None) == -1  # dt must be positive
    assert h.trot_qr_norm(d, 4, 300, 257, d, d, big, None) == -2
    assert h.trot_qr_norm(d, 4, 8, 2, d, d, 0, None) == -4
    assert h.trot_orthogonalize(None, 4, 8, 2, None, d, big, None) == -1
    # grouped launches: the descriptors are filled on the host, the launches refuse empty groups
    desc = ctypes.create_string_buffer(h.trot_gemm_problem_bytes())
    assert h.trot_gemm_problem(99, d, d, d, d, 4, 32, 8, 0.1, desc) == -1  # unknown operation
    assert h.trot_gemm_problem(1, d, d, d, None, 4, 32, 8, 0.1, desc) == -1  # Taylor needs the running sum
    assert h.trot_gemm_problem(0, d, d, d, None, 0, 32, 8, 0.0, desc) == -1  # no walkers
    wide = h.trot_gemm_problem(1, d, d, d, d, 4, 128, 16, 0.1, desc)
    narrow = h.trot_gemm_problem(5, d, d, d, d, 4, 128, 16, 0.1, desc)
    assert 0 < wide <= narrow  # the same problem in 32 x 64 and 32 x 32 tiles
    # the four tile shapes of the tuner as grouped Taylor problems: 128 / 64 or 32 row tiles x 16 / 32 or 64 columns
    assert [h.trot_gemm_problem(op, d, d, d, d, 4, 128, 16, 0.1, desc) for op in (7, 8, 1, 5)] == [8, 8, 16, 16]
    assert h.trot_gemm_problem(9, d, d, d, None, 4, 128, 16, 0.0, desc) == 8  # one-body on 64 x 32 tiles
    assert h.trot_gemm_problem(7, d, d, d, None, 4, 128, 16, 0.1, desc) == -1
    assert h.trot_gemm_f64_grouped(0, d, d, 0, 1, None) == -1
    assert h.trot_gemm_f64_grouped(99, d, d, 1, 1, None) == -1
    idesc = ctypes.create_string_buffer(h.trot_inv_problem_bytes())
    assert h.trot_inv_problem(d, d, d, 4, 65, idesc) == -2  # nocc above the kernels' limit
    assert h.trot_inv_problem(d, d, d, 5, 8, idesc) == 5  # one inverse per walker
    assert h.trot_inv_det_grouped(None, d, 1, 1, None) == -1
    cdesc = ctypes.create_string_buffer(h.trot_copy_problem_bytes())
    assert h.trot_copy_problem(d, d, 17, cdesc) == -1  # sizes are multiples of 16 bytes
    assert h.trot_copy_problem(d, d, 32, cdesc) == 0
    assert h.trot_copy_grouped(d, 0, 32, None) == -1
    assert h.trot_force_bias_slices(128, 800, 8000) == 31  # ~2048 workgroups, slices of >= 256, <= 64 slices
    assert h.trot_force_bias_slices(2048, 2560, 32768) == 2
    assert h.trot_force_bias_slices(4096, 4096, 32768) == 1
    assert h.trot_force_bias_slices(0, 8, 8) == -1
    gdesc = ctypes.create_string_buffer(h.trot_gemm_problem_bytes())
    assert h.trot_force_bias_problem(d, d, 128, 800, 8000, 1.0, 0, d, gdesc) == 26 * 31  # tiles x slices
    assert h.trot_force_bias_problem(None, d, 128, 800, 8000, 1.0, 0, d, gdesc) == -1
    sdesc = ctypes.create_string_buffer(h.trot_sum_problem_bytes())
    assert h.trot_sum_problem(d, d, 0, 100, 1.0, 0, sdesc) == -1
    assert h.trot_sum_problem(d, d, 3, 1000, 1.0, 0, sdesc) == 4  # blocks of 256 elements
    assert h.trot_sum_grouped(d, d, 0, 1, None) == -1
    hdesc = ctypes.create_string_buffer(h.trot_shift_problem_bytes())
    assert h.trot_shift_problem(d, d, d, 0.1, 0, 8, d, d, d, hdesc) == -1
    assert h.trot_shift_problem(d, d, d, 0.1, 5, 8, d, d, d, hdesc) == 5  # one workgroup per walker
    assert h.trot_shift_grouped(None, d, 1, 1, None) == -1
    udesc = ctypes.create_string_buffer(h.trot_update_problem_bytes())
    assert h.trot_update_problem(d, None, d, d, d, d, d, d, 4, 0.0, 0.0, 0.0, 1e-3, 100.0, 0.1, udesc) == -1
    assert h.trot_update_problem(d, None, d, d, d, d, d, d, 4, 0.01, 0.0, 0.0, 1e-3, 100.0, 0.1, udesc) == 0
    assert h.trot_update_grouped(d, 0, None) == -1

    # block observables (error paths only)
    assert h.trot_rdm1_from_green(None, d, 4, 8, 2, 2, 0, d, None) == -1
    assert h.trot_rdm1_from_green(d, d, 4, 8, 2, 2, 2, d, None) == -1  # spin slot out of range
    assert h.trot_rdm1_from_green(d, d, 4, 8, 9, 2, 0, d, None) == -1  # nocc > norb
    assert h.trot_rdm1_from_green(d, d, 4, 80, 65, 2, 0, d, None) == -2  # nocc > 64
    assert h.trot_density_corr(d, 0, 8, d, None) == -1
    assert h.trot_weighted_mean(d, d, d, 4, 10, d, d, 0, None) == -1  # workspace too small
    assert h.trot_weighted_mean(d, d, None, 4, 10, d, d, 1 << 20, None) == -1


def test_workspace_queries_are_sizes():
    from trot_amd import _lib

    h = _lib.lib()
    W, N, n = 16, 32, 4
    assert h.trot_taylor_workspace(W, N, n) >= 2 * W * N * n * 16
    assert h.trot_orthonormalize_workspace(W, N, n) >= W * N * n * 16 + 2 * W * n * n * 16
    assert h.trot_apply_trotter_workspace(W, N, n, 4) < h.trot_apply_trotter_workspace(W, N, n, 0)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from trot_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "absent.so")
    with pytest.raises(ImportError):
        _lib.lib()
