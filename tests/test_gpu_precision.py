"""Reduced-precision forms (FP32 and 3-pass split-BF16 MFMA) against the FP64 oracle.

Tolerances (max abs error / max abs value): FP32 GEMMs 2e-6; split-BF16
2e-4.  A single-pass BF16 product would miss these by about 10x (relative
error ~ 2^-8), so the BF16 bound also checks that all three passes are used.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import afqmc_oracle as O
from tests.cases import problem, rel, rel_each

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-6, "bf16x3": 2e-4}


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
@pytest.mark.parametrize("shape", [(64, 16, 96, 64), (37, 6, 70, 5), (24, 13, 33, 130)])
def test_vhs_and_taylor_lp(gpu, prec, shape):
    from trot_amd import _lib

    N, n, G, W = shape
    p = problem(N, n, G, W, seed=31)
    rng = np.random.default_rng(1)
    xbar = rng.standard_normal((W, G)) + 1j * rng.standard_normal((W, G))
    chol_flat = p["chol"].reshape(G, -1)
    v_ref = O.make_vhs(xbar, chol_flat, N)
    c32 = _t(chol_flat).float()
    v = _lib.vhs_build_lp(_t(xbar), c32, prec).reshape(W, N, N)
    assert v.dtype == torch.complex64
    assert rel(v.cpu().numpy(), v_ref) < TOL[prec]
    v64 = _lib.vhs_build_lp(_t(xbar), c32, prec, out_f64=True).reshape(W, N, N)
    assert v64.dtype == torch.complex128 and rel(v64.cpu().numpy(), v_ref) < TOL[prec]
    sdt = 0.1
    ref = O.taylor_expm_action(1j * sdt, v_ref, p["wa"], 6)
    out = _lib.taylor_apply_lp(v.contiguous(), _t(p["wa"]), sdt, 6, prec)
    # the identity term is exact; compare the propagated change against its own size
    d_ref = ref - p["wa"]
    assert rel(out.cpu().numpy() - p["wa"], d_ref) < 4 * TOL[prec]


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
@pytest.mark.parametrize("kind", ["rr", "ru", "uu"])
def test_force_bias_and_energy_lp(gpu, prec, kind):
    from trot_amd import meas, trial
    from trot_amd.core import HamChol, System

    N, n, G, W = 40, 8, 50, 12
    nb = 7 if kind == "uu" else n
    p = problem(N, n, G, W, seed=32, nocc_b=nb)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    if kind == "uu":
        sys_ = System(N, (n, nb), "unrestricted")
        tr = trial.UhfTrial(p["ca"], p["cb"])
        ops = meas.make_uhf_meas_ops(sys_, precision=prec)
        ref = O.make_uhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], p["cb"])
        w, wn = (_t(p["wa"]), _t(p["wb"])), (p["wa"], p["wb"])
    elif kind == "ru":
        sys_ = System(N, (n, n), "unrestricted")
        tr = trial.RhfTrial(p["ca"])
        ops = meas.make_rhf_meas_ops(sys_, precision=prec)
        ref = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], "unrestricted")
        w, wn = (_t(p["wa"]), _t(p["wb"])), (p["wa"], p["wb"])
    else:
        sys_ = System(N, (n, n), "restricted")
        tr = trial.RhfTrial(p["ca"])
        ops = meas.make_rhf_meas_ops(sys_, precision=prec)
        ref = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"])
        w, wn = _t(p["wa"]), p["wa"]
    ctx = ops.build_meas_ctx(ham, tr)
    fb = ops.require_kernel("force_bias")(w, ham, ctx, tr).cpu().numpy()
    e = ops.require_kernel("energy")(w, ham, ctx, tr).cpu().numpy()
    assert rel(fb, ref.force_bias(wn)) < TOL[prec]
    e_ref = ref.energy(wn)
    assert rel(e - p["h0"], e_ref - p["h0"]) < 4 * TOL[prec]


@pytest.mark.parametrize("shape", [(128, 128, 64, 1), (130, 70, 33, 1), (5, 300, 1000, 4), (257, 129, 96, 3)])
@pytest.mark.parametrize("cx", [True, False])
def test_split_bf16_nt_gemm(gpu, shape, cx):
    """trot_gemm_b3 against a float64 product, ragged sizes and split K."""
    from trot_amd import _b3

    M, N, K, ks = shape
    rng = np.random.default_rng(M + N + K)
    a = rng.standard_normal((M, K)) + (1j * rng.standard_normal((M, K)) if cx else 0)
    b = rng.standard_normal((N, K))
    ref = 0.7 * (a @ b.T)
    ap, bp = _b3.split(_t(a)), _b3.split(_t(b.T.copy()), transposed=True)
    for dt in ((torch.complex128, torch.complex64) if cx else (torch.float64, torch.float32)):
        if ks > 1 and dt != torch.complex128:
            continue
        out = torch.zeros((M, N), dtype=dt, device="cuda")
        _b3.gemm_nt(ap, bp, out, alpha=0.7, ksplit=ks)
        assert rel(out.cpu().numpy(), ref) < TOL["bf16x3"]
    if cx:  # beta accumulation and complex alpha
        out = torch.from_numpy(ref.copy()).cuda()
        _b3.gemm_nt(ap, bp, out, alpha=0.5j, beta=1.0, ksplit=ks)
        assert rel(out.cpu().numpy(), ref + 0.5j * (a @ b.T)) < TOL["bf16x3"]
        # split complex64 output (hi and lo words of bf16 re | bf16 im << 16); N = 70, 129 take the
        # element-by-element stores, 128 and 300 the 16-byte ones
        words = _b3.gemm_nt_split(ap, bp).cpu().numpy().view(np.uint32)

        def bf(u):
            return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)

        v = bf(words & 0xFFFF) + bf(words >> 16) * 1j  # each plane as complex
        assert rel(v[0] + v[1], a @ b.T) < TOL["bf16x3"]
        assert np.abs(v[1]).max() <= 2.0**-8 * np.abs(v[0]).max()


B3_SHAPES = [(64, 16, 30, 7), (37, 5, 21, 3), (130, 64, 40, 4), (512, 64, 16, 2), (96, 70, 20, 9)]


def _b3_case(shape, seed):
    from trot_amd import _lib

    N, n, G, W = shape
    p = problem(N, n, G, W, seed=seed)
    rng = np.random.default_rng(seed + 1)
    xbar = rng.standard_normal((W, G)) + 1j * rng.standard_normal((W, G))
    v_ref = O.make_vhs(xbar, p["chol"].reshape(G, -1), N)
    blk32 = _lib.pack_symblk(_t(p["chol"]), out_f32=True)
    v32 = _lib.vhs_build_lp(_t(xbar), blk32, "f32")
    w = rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n))
    return p, v_ref, v32, w


@pytest.mark.parametrize("shape", B3_SHAPES)
@pytest.mark.parametrize("n_terms", [1, 2, 6])
def test_taylor_b3_symblk(gpu, shape, n_terms):
    """Split-BF16 Taylor (V read from sym-block complex64, FP32 term sum) against the oracle.

    Ragged N (37, 130), n above one 64-column tile (70), walker counts that are not a
    multiple of 8, and 1 / 2 / 6 terms (first-and-last, first-then-last, middle terms).
    """
    from trot_amd import _lib

    _, v_ref, v32, w = _b3_case(shape, 35)
    sdt = 0.1
    ref = O.taylor_expm_action(1j * sdt, v_ref, w, n_terms)
    out = _lib.taylor_apply_b3_symblk(v32, _t(w), sdt, n_terms)
    # the identity term is exact; compare the propagated change against its own size
    assert rel(out.cpu().numpy() - w, ref - w) < 4 * TOL["bf16x3"]
    out0 = _lib.taylor_apply_b3_symblk(v32, _t(w), sdt, 0)
    assert np.array_equal(out0.cpu().numpy(), w)


@pytest.mark.parametrize("shape", B3_SHAPES)
@pytest.mark.parametrize("n_terms", [0, 1, 6])
def test_trotter_b3_symblk(gpu, shape, n_terms):
    """E . taylor(E w) on split-BF16 MFMA against the oracle's FP64 products, in and out of place."""
    import scipy.linalg
    from trot_amd import _lib

    N = shape[0]
    p, v_ref, v32, w = _b3_case(shape, 36)
    E = scipy.linalg.expm(-0.05 * p["h1"])
    sdt = 0.1
    ref = np.einsum("ij,wjp->wip", E, O.taylor_expm_action(1j * sdt, v_ref, np.einsum("ij,wjp->wip", E, w), n_terms))
    planes = _lib.onebody_b3_planes(_t(E))
    wt = _t(w)
    out = _lib.trotter_b3_symblk(v32, wt, planes, sdt, n_terms)
    assert np.array_equal(wt.cpu().numpy(), w)  # input untouched out of place
    assert rel(out.cpu().numpy() - w, ref - w) < 4 * TOL["bf16x3"]
    out2 = _lib.trotter_b3_symblk(v32, wt, planes, sdt, n_terms, out=wt)
    assert out2.data_ptr() == wt.data_ptr()
    assert np.array_equal(out2.cpu().numpy(), out.cpu().numpy())
    with pytest.raises(ValueError):
        _lib.trotter_b3_symblk(v32, wt, _lib.onebody_b3_planes(_t(np.eye(N + 128))), sdt, n_terms)


def _bf16_split_host(x: np.ndarray):
    """hi = bf16(x), lo = bf16(x - hi), round to nearest even, as 16-bit words (x float32, finite)."""
    def rne(v):
        u = v.astype(np.float32).view(np.uint32).astype(np.uint64)
        return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)

    hi = rne(x)
    lo = rne(x - (hi << 16).view(np.float32))
    return hi, lo


def _split_v_case(shape, seed):
    from trot_amd import _b3, _lib

    N, n, G, W = shape
    p = problem(N, n, G, W, seed=seed)
    rng = np.random.default_rng(seed + 1)
    xbar = rng.standard_normal((W, G)) + 1j * rng.standard_normal((W, G))
    chol_planes = _b3.split(_lib.pack_symblk(_t(p["chol"])), transposed=True)
    xp = _b3.split(_t(xbar))
    v64 = torch.empty((W, chol_planes.rows), dtype=torch.complex64, device="cuda")
    _b3.gemm_nt(xp, chol_planes, v64)
    vs = _b3.gemm_nt_split(xp, chol_planes)
    v_ref = O.make_vhs(xbar, p["chol"].reshape(G, -1), N)
    w = rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n))
    return p, v_ref, v64, vs, w


@pytest.mark.parametrize("shape", [(64, 16, 30, 7), (37, 5, 21, 3), (512, 64, 40, 9)])
def test_split_v_planes_and_propagation(gpu, shape):
    """VHS written as split planes: the planes are the bf16 hi / lo split of the complex64 VHS bit for bit;
    with every term on both planes the propagation equals the complex64 form bit for bit; with terms
    k >= 3 as one bf16 pass, the result moves by no more than 2^-7 of the size of those terms."""
    import scipy.linalg
    from trot_amd import _lib

    N = shape[0]
    p, v_ref, v64, vs, w = _split_v_case(shape, 41)
    vc = torch.view_as_real(v64).cpu().numpy()
    words = vs.cpu().numpy().view(np.uint32)
    re_hi, re_lo = _bf16_split_host(vc[..., 0])
    im_hi, im_lo = _bf16_split_host(vc[..., 1])
    assert np.array_equal(words[0] & 0xFFFF, re_hi) and np.array_equal(words[0] >> 16, im_hi)
    assert np.array_equal(words[1] & 0xFFFF, re_lo) and np.array_equal(words[1] >> 16, im_lo)

    E = scipy.linalg.expm(-0.05 * p["h1"])
    planes = _lib.onebody_b3_planes(_t(E))
    sdt, K = 0.1, 6
    wt = _t(w)
    full = _lib.trotter_b3_symblk(v64, wt, planes, sdt, K)
    same = _lib.trotter_b3v(vs, wt, planes, sdt, K, K + 1)
    assert torch.equal(full, same)
    hi3 = _lib.trotter_b3v(vs, wt, planes, sdt, K, 3)
    assert not torch.equal(full, hi3)
    ew = np.einsum("ij,wjp->wip", E, w)
    late = np.einsum("ij,wjp->wip", E, O.taylor_expm_action(1j * sdt, v_ref, ew, K)
                     - O.taylor_expm_action(1j * sdt, v_ref, ew, 2))  # E (T_3 + ... + T_6)
    d = np.abs(hi3.cpu().numpy() - full.cpu().numpy()).max()
    assert d <= 2.0**-7 * np.abs(late).max()
    ref = np.einsum("ij,wjp->wip", E, O.taylor_expm_action(1j * sdt, v_ref, ew, K))
    assert rel(hi3.cpu().numpy() - w, ref - w) < 4 * TOL["bf16x3"]


@pytest.mark.parametrize("prec,tprec", [("f32", "f32"), ("bf16x3", "f32"), ("bf16x3", "bf16x3"), ("f32", "f64")])
def test_mixed_steps_track_fp64(gpu, prec, tprec):
    """A few mixed-precision steps stay close to the FP64 oracle with identical fields."""
    from trot_amd import meas, prop, trial
    from trot_amd.core import HamChol, System

    N, n, G, W = 48, 8, 60, 32
    p = problem(N, n, G, W, seed=33)
    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    mops = meas.make_rhf_meas_ops(sys_, precision=prec)
    tops = trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted", precision=prec, taylor_precision=tprec)
    params = prop.QmcParams(dt=0.01, n_walkers=W, seed=1)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=_t(p["wa"]))
    ref_ops = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"])
    ref_ctx = O.build_prop_ctx(p["h0"], p["h1"], p["chol"], O.rhf_rdm1(p["ca"]), 0.01)
    ref = O.init_prop_state(p["wa"], ref_ops, W)
    for s in range(4):
        f = p["fields"][s]
        st = pops.step(st, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
                       prop_ctx=pctx, fields=_t(f))
        ref = O.afqmc_step(ref, f, ref_ctx, ref_ops, O.QmcParams(dt=0.01, n_walkers=W))
    tol = 50 * TOL[prec if tprec != "bf16x3" else "bf16x3"]
    assert rel(st.walkers.cpu().numpy(), ref.walkers) < tol
    assert rel_each(st.weights.cpu().numpy(), ref.weights) < tol


@pytest.mark.parametrize("shape", [(40, 8, 50, 13), (130, 64, 7, 4), (64, 1, 3, 1), (200, 33, 130, 6)])
@pytest.mark.parametrize("kind", ["rr", "uu"])
def test_local_energy_b3(gpu, shape, kind):
    """Batched split-BF16 exchange GEMM energy against the FP64 oracle and the FP64 kernel.

    Odd walker counts (the last 2-walker tile half empty), G not a multiple of 4, nocc = 64 (full
    64-row blocks) and nocc = 1.
    """
    from trot_amd import _lib, meas, trial
    from trot_amd.core import HamChol, System

    N, n, G, W = shape
    nb = max(1, n - 1) if kind == "uu" else n
    p = problem(N, n, G, W, seed=37, nocc_b=nb)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    if kind == "uu":
        sys_ = System(N, (n, nb), "unrestricted")
        tr = trial.UhfTrial(p["ca"], p["cb"])
        ops = meas.make_uhf_meas_ops(sys_, precision="bf16x3")
        ops64 = meas.make_uhf_meas_ops(sys_)
        ref = O.make_uhf_ops(p["h0"], p["h1"], p["chol"], p["ca"], p["cb"])
        w, wn = (_t(p["wa"]), _t(p["wb"])), (p["wa"], p["wb"])
    else:
        sys_ = System(N, (n, n), "restricted")
        tr = trial.RhfTrial(p["ca"])
        ops = meas.make_rhf_meas_ops(sys_, precision="bf16x3")
        ops64 = meas.make_rhf_meas_ops(sys_)
        ref = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"])
        w, wn = _t(p["wa"]), p["wa"]
    e = ops.require_kernel("energy")(w, ham, ops.build_meas_ctx(ham, tr), tr).cpu().numpy()
    e64 = ops64.require_kernel("energy")(w, ham, ops64.build_meas_ctx(ham, tr), tr).cpu().numpy()
    e_ref = ref.energy(wn)
    assert rel(e64 - p["h0"], e_ref - p["h0"]) < 1e-10
    assert rel(e - p["h0"], e_ref - p["h0"]) < 4 * TOL["bf16x3"]


def test_reference_mixed_precision_flag(gpu):
    """make_prop_ops(mixed_precision=True) is the reference's mixed mode: xbar as complex64 and chol_flat as
    float32 for the VHS only, V cast back to complex128, one-body factors and Taylor series in FP64
    (trot/prop/chol_afqmc_ops.py:160, 247-257; trot/prop/afqmc.py:177).

    Part 1: with xbar and every L_g multiples of 1/64 below 2 in magnitude and G = 16, every product and
    partial sum of the VHS is exact in FP32, so the FP32 VHS equals the FP64 one and the whole Trotter factor
    must meet the FP64 oracle at 1e-12; an FP32 one-body or Taylor product would miss that by ~1e5.
    Part 2: with general inputs, four steps stay within the FP32 bound of the oracle run with the reference's
    casts (float32 chol_flat).
    """
    from dataclasses import replace

    from trot_amd import meas, prop, trial
    from trot_amd.core import HamChol, System

    N, n, G, W = 40, 6, 16, 9
    rng = np.random.default_rng(77)
    a = rng.standard_normal((N, N)) / np.sqrt(N)
    h1 = np.triu(a) + np.triu(a, 1).T
    b = np.round(rng.uniform(-1.9, 1.9, (G, N, N)) * 64) / 64
    chol = np.triu(b) + np.swapaxes(np.triu(b, 1), 1, 2)  # symmetric, still multiples of 1/64
    chol *= 1.0 / 64  # |L| < 2/64: the products stay multiples of 2^-18, sums below 2^6
    ca = np.linalg.eigh(h1)[1][:, :n]
    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(0.0, h1, chol).to_device()
    tr = trial.RhfTrial(ca)
    tops = trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted", mixed_precision=True)
    params = prop.QmcParams(dt=0.01, n_walkers=W, seed=1)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    assert pctx.precision == "f32" and pctx.taylor_precision == "f64"
    ref_ctx = O.build_prop_ctx(0.0, h1, chol, O.rhf_rdm1(ca), 0.01)
    w = ca[None] + 0.05 * (rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n)))
    xbar = (np.round(rng.uniform(-1.9, 1.9, (W, G)) * 64) + 1j * np.round(rng.uniform(-1.9, 1.9, (W, G)) * 64)) / 64
    got = prop.apply_trotter(_t(w), _t(xbar), pctx, 6).cpu().numpy()
    exact = O.apply_trotter_r(w, xbar, ref_ctx, 6)
    assert rel(got - w, exact - w) < 1e-12

    # part 2: general inputs against the oracle with the reference's float32 chol_flat
    p = problem(48, 8, 60, 32, seed=33)
    ham = HamChol(p["h0"], p["h1"], p["chol"]).to_device()
    tr = trial.RhfTrial(p["ca"])
    sys_ = System(48, (8, 8), "restricted")
    mops = meas.make_rhf_meas_ops(sys_)
    tops = trial.make_rhf_trial_ops(sys_)
    params = prop.QmcParams(dt=0.01, n_walkers=32, seed=1)
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=_t(p["wa"]))
    ref_ops = O.make_rhf_ops(p["h0"], p["h1"], p["chol"], p["ca"])
    ref64_ctx = O.build_prop_ctx(p["h0"], p["h1"], p["chol"], O.rhf_rdm1(p["ca"]), 0.01)
    ref_ctx = replace(ref64_ctx, chol_flat=ref64_ctx.chol_flat.astype(np.float32))
    ref = O.init_prop_state(p["wa"], ref_ops, 32)
    ref64 = O.init_prop_state(p["wa"], ref_ops, 32)
    for s in range(4):
        f = p["fields"][s]
        st = pops.step(st, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
                       prop_ctx=pctx, fields=_t(f))
        ref = O.afqmc_step(ref, f, ref_ctx, ref_ops, O.QmcParams(dt=0.01, n_walkers=32))
        ref64 = O.afqmc_step(ref64, f, ref64_ctx, ref_ops, O.QmcParams(dt=0.01, n_walkers=32))
    wg = st.walkers.cpu().numpy()
    # the FP32 VHS error, relative to the walkers' change over four steps
    assert rel(wg - p["wa"], ref.walkers - p["wa"]) < 50 * TOL["f32"]
    assert rel_each(st.weights.cpu().numpy(), ref.weights) < 50 * TOL["f32"]
    # and the reference's mixed mode itself differs from FP64 by FP32 rounding only
    assert rel(ref.walkers - p["wa"], ref64.walkers - p["wa"]) < 50 * TOL["f32"]
