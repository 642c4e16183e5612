"""FP64 HIP kernels against the CPU oracle, one C entry point at a time (tolerance 1e-10 relative)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import afqmc_oracle as O
from tests.cases import problem, rel, rel_each

pytestmark = pytest.mark.gpu

TOL = 1e-10
# (norb, nocc, nchol, walkers): tile-aligned, ragged, tiny, single-occupied, single walker
SHAPES = [(64, 16, 64, 64), (37, 6, 70, 5), (24, 13, 33, 130), (9, 1, 5, 3), (50, 64 // 2, 17, 1), (70, 64, 20, 4)]


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("shape", SHAPES)
def test_overlap_green(gpu, shape):
    from trot_amd import _lib

    N, n, G, W = shape
    p = problem(N, n, G, W, seed=11)
    det, gt = _lib.overlap_green(_t(p["ca"]), _t(p["wa"]))
    ov = _lib.overlap_combine(det, None)
    ref_ov = O.overlap_r(p["wa"], p["ca"])
    ref_g = O.half_green(p["wa"], p["ca"])  # (W, n, N)
    torch.cuda.synchronize()
    assert rel_each(ov.cpu().numpy(), ref_ov) < TOL
    assert rel(gt.cpu().numpy(), np.swapaxes(ref_g, 1, 2)) < TOL


@pytest.mark.parametrize("shape", [(64, 64, 1, 9), (80, 17, 1, 6), (40, 33, 1, 3), (5, 1, 1, 2)])
def test_overlap_green_pivoting(gpu, shape):
    """Walkers far from the trial: C^T w has no dominant diagonal, so most elimination steps swap rows."""
    from trot_amd import _lib

    N, n, G, W = shape
    p = problem(N, n, G, W, seed=12)
    rng = np.random.default_rng(5)
    w = rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n))
    m = np.einsum("ip,wiq->wpq", p["ca"], w)
    # row-permuted identity-like blocks make the first pivots fall off the diagonal
    if n > 1:
        assert np.any(np.argmax(np.abs(m[:, :, 0].real) + np.abs(m[:, :, 0].imag), axis=1) != 0)
    det, gt = _lib.overlap_green(_t(p["ca"]), _t(w))
    torch.cuda.synchronize()
    assert rel(det.cpu().numpy(), np.linalg.det(m)) < TOL
    assert rel(gt.cpu().numpy(), np.swapaxes(O.half_green(w, p["ca"]), 1, 2)) < TOL


@pytest.mark.parametrize("shape", [(70, 11, 40, 7), (200, 40, 97, 128)])
def test_force_bias_split_k(gpu, shape):
    """K = N n split into slices (uneven last slice), FP64 partials summed in order; accumulate path."""
    from trot_amd import _lib

    N, n, G, W = shape
    lib = _lib.lib()
    assert lib.trot_force_bias_workspace(W, G, N * n) > 0  # the split path runs
    p = problem(N, n, G, W, seed=13)
    rc = np.einsum("ip,gij->gpj", p["ca"], p["chol"])  # rot_chol (G, n, N)
    ref = 2.0 * np.einsum("gpj,wpj->wg", rc, O.half_green(p["wa"], p["ca"]))  # trot/meas/rhf.py:119-125
    _, gt = _lib.overlap_green(_t(p["ca"]), _t(p["wa"]))
    rct = _t(np.ascontiguousarray(np.swapaxes(rc, 1, 2)))
    fb = _lib.force_bias(gt, rct, 2.0)
    assert rel(fb.cpu().numpy(), ref) < TOL
    fb2 = _lib.force_bias(gt, rct, 1.0, out=fb.clone(), accumulate=True)
    assert rel(fb2.cpu().numpy(), 1.5 * ref) < TOL


@pytest.mark.parametrize("shape", SHAPES[:4])
def test_force_bias_restricted(gpu, shape):
    from trot_amd import meas, trial
    from trot_amd.core import HamChol, System

    N, n, G, W = shape
    p = problem(N, n, G, W, seed=12)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    ops = meas.make_rhf_meas_ops(System(N, (n, n), "restricted"))
    ctx = ops.build_meas_ctx(ham, tr)
    fb = ops.require_kernel("force_bias")(_t(p["wa"]), ham, ctx, tr).cpu().numpy()
    ref_ctx = O.build_meas_ctx_rhf(p["h1"], p["chol"], p["ca"])
    assert rel(ctx.rot_chol.cpu().numpy(), ref_ctx.rot_chol) < TOL
    assert rel(ctx.rot_h1.cpu().numpy(), ref_ctx.rot_h1) < TOL
    assert rel(fb, O.force_bias_rw(p["wa"], ref_ctx, p["ca"])) < TOL


def test_force_bias_and_energy_unrestricted_rhf_trial(gpu):
    from trot_amd import meas, trial
    from trot_amd.core import HamChol, System

    N, n, G, W = 30, 5, 40, 9
    p = problem(N, n, G, W, seed=13, nocc_b=5)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    ops = meas.make_rhf_meas_ops(System(N, (n, n), "unrestricted"))
    ctx = ops.build_meas_ctx(ham, tr)
    w = (_t(p["wa"]), _t(p["wb"]))
    fb = ops.require_kernel("force_bias")(w, ham, ctx, tr).cpu().numpy()
    e = ops.require_kernel("energy")(w, ham, ctx, tr).cpu().numpy()
    ref = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], "unrestricted")
    assert rel(fb, ref.force_bias((p["wa"], p["wb"]))) < TOL
    assert rel(e, ref.energy((p["wa"], p["wb"]))) < TOL


@pytest.mark.parametrize("nab", [(7, 6), (12, 11), (1, 1)])
def test_uhf_force_bias_energy_overlap(gpu, nab):
    from trot_amd import meas, trial
    from trot_amd.core import HamChol, System

    na, nb = nab
    N, G, W = 40, 31, 7
    p = problem(N, na, G, W, seed=14, nocc_b=nb)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.UhfTrial(p["ca"], p["cb"])
    ops = meas.make_uhf_meas_ops(System(N, (na, nb), "unrestricted"))
    ctx = ops.build_meas_ctx(ham, tr)
    w = (_t(p["wa"]), _t(p["wb"]))
    ref = O.make_uhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], p["cb"])
    wn = (p["wa"], p["wb"])
    assert rel_each(ops.overlap(w, tr).cpu().numpy(), ref.overlap(wn)) < TOL
    assert rel(ops.require_kernel("force_bias")(w, ham, ctx, tr).cpu().numpy(), ref.force_bias(wn)) < TOL
    assert rel(ops.require_kernel("energy")(w, ham, ctx, tr).cpu().numpy(), ref.energy(wn)) < TOL


@pytest.mark.parametrize("shape", SHAPES)
def test_energy_restricted(gpu, shape):
    from trot_amd import meas, trial
    from trot_amd.core import HamChol, System

    N, n, G, W = shape
    p = problem(N, n, G, W, seed=15)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    ops = meas.make_rhf_meas_ops(System(N, (n, n), "restricted"))
    ctx = ops.build_meas_ctx(ham, tr)
    e = ops.require_kernel("energy")(_t(p["wa"]), ham, ctx, tr).cpu().numpy()
    ref = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"])
    assert rel(e, ref.energy(p["wa"])) < TOL


def test_energy_single_walker_view(gpu):
    """A single walker (no leading axis) is measured like a batch of one (per-walker kernel protocol)."""
    from trot_amd import meas, trial
    from trot_amd.core import HamChol, System

    p = problem(16, 4, 12, 3, seed=16)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    ops = meas.make_rhf_meas_ops(System(16, (4, 4), "restricted"))
    ctx = ops.build_meas_ctx(ham, tr)
    e1 = ops.require_kernel("energy")(_t(p["wa"][1]), ham, ctx, tr)
    assert e1.dim() == 0
    ref = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"]).energy(p["wa"][1:2])[0]
    assert abs(complex(e1.cpu()) - ref) / abs(ref) < TOL


@pytest.mark.parametrize("shape", SHAPES[:4])
def test_vhs_taylor_trotter(gpu, shape):
    from trot_amd import _lib

    N, n, G, W = shape
    p = problem(N, n, G, W, seed=17)
    rng = np.random.default_rng(5)
    xbar = rng.standard_normal((W, G)) + 1j * rng.standard_normal((W, G))
    chol_flat = p["chol"].reshape(G, -1)
    v = _lib.vhs_build(_t(xbar), _t(chol_flat)).reshape(W, N, N)
    v_ref = O.make_vhs(xbar, chol_flat, N)
    assert rel(v.cpu().numpy(), v_ref) < TOL
    sdt = np.sqrt(0.01)
    out = _lib.taylor_apply(v, _t(p["wa"]), sdt, 6)
    ref = O.taylor_expm_action(1j * sdt, v_ref, p["wa"], 6)
    assert rel(out.cpu().numpy(), ref) < TOL
    e = np.linalg.qr(rng.standard_normal((N, N)))[0] * 0.9
    ob = _lib.one_body_apply(_t(e), _t(p["wa"]))
    assert rel(ob.cpu().numpy(), e @ p["wa"]) < TOL
    ctx = O.PropCtx(dt=0.01, sqrt_dt=sdt, exp_h1_half=e, mf_shifts=np.zeros(G), h0_prop=0.0, chol_flat=chol_flat,
                    norb=N)
    for chunk in (0, 2):
        wa, _ = _lib.apply_trotter(_t(p["wa"]), None, _t(xbar), _t(chol_flat), _t(e), sdt, 6, chunk)
        assert rel(wa.cpu().numpy(), O.apply_trotter_r(p["wa"], xbar, ctx, 6)) < TOL


def test_field_shift(gpu):
    from trot_amd import _lib

    W, G = 37, 301
    rng = np.random.default_rng(3)
    x = rng.standard_normal((W, G))
    fb = rng.standard_normal((W, G)) + 1j * rng.standard_normal((W, G))
    mf = 1j * rng.standard_normal(G) + 0.1 * rng.standard_normal(G)
    ctx = O.PropCtx(dt=0.01, sqrt_dt=0.1, exp_h1_half=None, mf_shifts=mf, h0_prop=0.0, chol_flat=None, norb=0)
    xb, st, ft = _lib.field_shift(_t(x), _t(fb), _t(mf), 0.1)
    rxb, rst, rft = O.shift_terms(x, fb, ctx)
    assert rel(xb.cpu().numpy(), rxb) < TOL
    assert rel(st.cpu().numpy(), rst) < TOL
    assert rel(ft.cpu().numpy(), rft) < TOL


def test_phaseless_update_masking(gpu):
    """Every masking branch of trot/prop/afqmc.py:122-130 against the oracle."""
    from trot_amd import _lib

    W = 1500
    rng = np.random.default_rng(4)
    ov_old = rng.standard_normal(W) + 1j * rng.standard_normal(W)
    ov_new = ov_old * (1 + 0.1 * (rng.standard_normal(W) + 1j * rng.standard_normal(W)))
    ov_new[0] = -ov_old[0]  # phase pi: importance < 0 -> node
    ov_new[1] = 1j * ov_old[1]  # cos(theta) = 0
    ov_new[2] = np.nan  # non-finite
    ov_new[3] = ov_old[3] * 1e5  # importance above cap
    ov_new[4] = ov_old[4] * 1e-5  # below floor
    weights = np.abs(rng.standard_normal(W)) + 0.5
    weights[5] = 99.99  # weight above cap after multiplication
    ov_new[5] = ov_old[5] * 1.5
    st = 0.1 * (rng.standard_normal(W) + 1j * rng.standard_normal(W))
    ft = 0.1 * (rng.standard_normal(W) + 1j * rng.standard_normal(W))
    params = O.QmcParams(dt=0.01)
    ctx = O.PropCtx(dt=0.01, sqrt_dt=0.1, exp_h1_half=None, mf_shifts=None, h0_prop=complex(-1.3, 0.0),
                    chol_flat=None, norb=0)
    w_ref, nodes_ref, pop_ref = O.phaseless_update(weights, ov_old, ov_new, st, ft, 0.2, -1.1, 3, ctx, params)
    w = _t(weights.copy())
    scal = _t(np.array([0.2, -1.1]))
    nodes = torch.tensor([3], dtype=torch.int64, device="cuda")
    _lib.phaseless_update(w, _t(ov_old), _t(ov_new), _t(st), _t(ft), scal, nodes, 0.01, complex(-1.3, 0.0), 1e-3,
                          100.0, 0.1)
    assert rel(w.cpu().numpy(), w_ref) < TOL
    assert int(nodes.item()) == nodes_ref
    assert abs(scal[0].item() - pop_ref) < 1e-10 * abs(pop_ref)
    assert w.cpu().numpy()[[0, 1, 2, 3, 4, 5]].tolist() == [0.0] * 6


@pytest.mark.parametrize("shape", [(64, 16, 1, 64), (37, 6, 1, 5), (200, 40, 1, 33), (12, 12, 1, 2)])
def test_orthonormalize(gpu, shape):
    from trot_amd import _lib

    N, n, _, W = shape
    rng = np.random.default_rng(6)
    w = rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n))
    q = _lib.orthonormalize(_t(w)).cpu().numpy()
    q_ref, _ = O.qr_phase(w)
    assert rel(q, q_ref) < 1e-10


@pytest.mark.parametrize("W", [1, 7, 1024, 3000])
def test_sr_comb_and_gather(gpu, W):
    from trot_amd import _lib

    rng = np.random.default_rng(W)
    weights = np.abs(rng.standard_normal(W))
    weights[::5] = 0.0
    if W == 1:
        weights[:] = 0.7
    zeta = 0.3141
    idx, wn = _lib.sr_comb(_t(weights), zeta)
    ref_idx = O.sr_indices(weights, zeta, W)
    assert np.array_equal(idx.cpu().numpy(), np.minimum(ref_idx, W - 1))
    assert np.allclose(wn.cpu().numpy(), np.sum(weights) / W, rtol=1e-14)
    src = rng.standard_normal((W, 3, 5)) + 1j * rng.standard_normal((W, 3, 5))
    g = _lib.gather_rows(_t(src), idx).cpu().numpy()
    assert np.array_equal(g, src[idx.cpu().numpy()])


def test_prop_ctx(gpu):
    from trot_amd import prop, trial
    from trot_amd.core import HamChol

    N, n, G = 33, 7, 45
    p = problem(N, n, G, 1, seed=18)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    ctx = prop.build_prop_ctx(ham, trial.rhf_get_rdm1(tr), 0.01)
    ref = O.build_prop_ctx(p["h0"], p["h1"], p["chol"], O.rhf_rdm1(p["ca"]), 0.01)
    assert rel(ctx.mf_shifts.cpu().numpy(), ref.mf_shifts) < TOL
    assert abs(ctx.h0_prop - ref.h0_prop) < 1e-12 * abs(ref.h0_prop)
    assert rel(ctx.exp_h1_half.cpu().numpy(), ref.exp_h1_half) < TOL


def test_bad_sizes_raise(gpu):
    from trot_amd import _lib

    w = torch.zeros((2, 300, 257), dtype=torch.complex128, device="cuda")  # nocc above 256
    c = torch.zeros((300, 257), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        _lib.overlap_green(c, w)
    with pytest.raises(ValueError):
        _lib.overlap_green(c[:, :3].contiguous(), w)


@pytest.mark.parametrize("N", [64, 37, 9])
def test_symblk_vhs_and_taylor(gpu, N):
    """Sym-block V (upper 8x8 blocks only) gives the dense VHS / Taylor results."""
    from trot_amd import _lib

    n, G, W = 5, 21, 6
    p = problem(N, n, G, W, seed=41)
    rng = np.random.default_rng(2)
    xbar = rng.standard_normal((W, G)) + 1j * rng.standard_normal((W, G))
    chol_flat = p["chol"].reshape(G, -1)
    blk = _lib.pack_symblk(_t(p["chol"]))
    nb = (N + 7) // 8
    assert blk.shape == (G, 64 * nb * (nb + 1) // 2)
    vpk = _lib.vhs_build(_t(xbar), blk)
    v_ref = O.make_vhs(xbar, chol_flat, N)
    # unpack on the host and compare with the dense VHS
    vp = vpk.cpu().numpy()
    dense = np.zeros((W, nb * 8, nb * 8), dtype=complex)
    pidx = 0
    for bi in range(nb):
        for bj in range(bi, nb):
            b = vp[:, pidx * 64:(pidx + 1) * 64].reshape(W, 8, 8)
            dense[:, bi * 8:bi * 8 + 8, bj * 8:bj * 8 + 8] = b
            dense[:, bj * 8:bj * 8 + 8, bi * 8:bi * 8 + 8] = np.swapaxes(b, 1, 2)
            pidx += 1
    assert rel(dense[:, :N, :N], v_ref) < TOL
    assert np.all(dense[:, N:, :] == 0) and np.all(dense[:, :, N:] == 0)
    sdt = 0.1
    ref = O.taylor_expm_action(1j * sdt, v_ref, p["wa"], 6)
    out = _lib.taylor_apply_symblk(vpk, _t(p["wa"]), sdt, 6)
    assert rel(out.cpu().numpy(), ref) < TOL
    for prec, tol in (("f32", 2e-6), ("bf16x3", 2e-4)):
        blk32 = _lib.pack_symblk(_t(p["chol"]), out_f32=True)
        v32 = _lib.vhs_build_lp(_t(xbar), blk32, prec)
        out = _lib.taylor_apply_lp_symblk(v32, _t(p["wa"]), sdt, 6, prec)
        assert rel(out.cpu().numpy() - p["wa"], ref - p["wa"]) < 4 * tol


@pytest.mark.parametrize("shape", [(40, 8, 6), (130, 33, 3), (64, 64, 2)])
@pytest.mark.parametrize("symblk", [False, True])
def test_taylor_tile_shapes_bit_identical(gpu, shape, symblk):
    """Every FP64 Taylor tile shape gives the same bits; the autotuner records one of them."""
    from trot_amd import _lib

    N, n, W = shape
    p = problem(N, n, 17, W, seed=19)
    rng = np.random.default_rng(3)
    xbar = rng.standard_normal((W, 17)) + 1j * rng.standard_normal((W, 17))
    v_ref = O.make_vhs(xbar, p["chol"].reshape(17, -1), N)
    ref = O.taylor_expm_action(1j * 0.1, v_ref, p["wa"], 6)
    if symblk:
        vin = _lib.vhs_build(_t(xbar), _lib.pack_symblk(_t(p["chol"])))
        run = lambda: _lib.taylor_apply_symblk(vin, _t(p["wa"]), 0.1, 6).cpu().numpy()  # noqa: E731
    else:
        vin = _t(v_ref)
        run = lambda: _lib.taylor_apply(vin, _t(p["wa"]), 0.1, 6).cpu().numpy()  # noqa: E731
    outs = []
    for tile in range(4):
        _lib.set_taylor_tile(W, N, n, symblk, tile)
        assert _lib.get_taylor_tile(W, N, n, symblk) == tile
        outs.append(run())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    assert rel(outs[0], ref) < TOL
    best, ms = _lib.autotune_taylor(W, N, n, symblk)
    assert 0 <= best < 4 and all(m > 0 for m in ms) and _lib.get_taylor_tile(W, N, n, symblk) == best
    assert np.array_equal(run(), outs[0])


def _conditioned_walkers(rng, N, n, kappas):
    """w = U diag(s) V^H with singular values from 1 down to 1/kappa (kappa = 1: orthonormal columns)."""
    out = []
    for kappa in kappas:
        u = np.linalg.qr(rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n)))[0]
        v = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
        s = np.logspace(0.0, -np.log10(kappa), n)
        out.append((u * s) @ v.conj().T)
    return np.array(out)


@pytest.mark.parametrize("N,n", [(40, 8), (130, 64)])
def test_orthonormalize_ill_conditioned(gpu, N, n):
    """Walkers with kappa 1e6 and 1e10 break Cholesky-QR (kappa^2 eps > 1); they are redone by Gram-Schmidt
    with reorthogonalisation, next to well-conditioned walkers that stay on Cholesky-QR.  The result is the
    reference's Q (trot/walkers.py:137-148, the oracle's qr_phase) to kappa eps, orthonormal to rounding,
    spans the walker's columns, and R = Q^H w is upper triangular with a real positive diagonal."""
    from trot_amd import _lib

    rng = np.random.default_rng(44)
    kappas = [1.0, 1e10, 3.0, 1e6, 1e10, 10.0]
    w = _conditioned_walkers(rng, N, n, kappas)
    q = _lib.orthonormalize(_t(w)).cpu().numpy()
    q_ref, _ = O.qr_phase(w)
    _, det_r = _lib.orthogonalize(_t(w))
    det_norm = _lib.qr_norm(_t(w)).cpu().numpy()
    for b, kappa in enumerate(kappas):
        assert np.all(np.isfinite(q[b])), b
        assert np.max(np.abs(q[b].conj().T @ q[b] - np.eye(n))) < 1e-12, b
        r = q[b].conj().T @ w[b]
        assert np.max(np.abs(w[b] - q[b] @ r)) < 1e-12 * np.max(np.abs(w[b])), b
        assert np.max(np.abs(np.tril(r, -1))) < 1e-12 * np.max(np.abs(w[b])), b
        d = np.diagonal(r)
        assert np.all(d.real > 0) and np.max(np.abs(d.imag)) < 1e-12 * np.max(np.abs(w[b])), b
        assert rel(q[b], q_ref[b]) < max(1e-10, 1e3 * kappa * 2.2e-16), b
        # |det R| = product of the singular values (where that product is a normal double)
        prod_s = np.prod(np.logspace(0.0, -np.log10(kappa), n))
        if prod_s > 1e-280:
            assert abs(det_r.cpu().numpy()[b] - prod_s) < max(1e-10, 1e3 * kappa * 2.2e-16) * prod_s, b
            assert abs(det_norm[b] - prod_s) < max(1e-10, 1e3 * kappa * 2.2e-16) * prod_s, b
