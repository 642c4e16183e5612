"""The differentiation form by reverse mode: dE_block/dh1 from the hand-written adjoint (trot_amd/adjoint.py).

* The building blocks: trot_gemm_batched for every (op(A), op(B)) with real and complex operands, strided
  outputs and shared operands against complex128 products on the CPU; trot_inverse_det against numpy.
* Small system (N = 24, n = 4, G = 30, 16 walkers, 12 steps): the adjoint's forward block energy equals the
  product's block path; <dE/dh1, D> for three random symmetric D against central differences of the block
  energy (1e-6 relative) and against the tangent propagation of trot_amd.diff (1e-8: two independent exact
  derivatives); single entries of dE/dh1 against central differences; the gradient does not depend on the
  checkpoint spacing (bit for bit); the FP32-contraction form within 1e-4 of the FP64 one; the autograd
  Function gives h1.grad and lam.grad.
* The same small system against torch autograd of the differentiable CPU restatement (oracle/autodiff.py,
  the stand-in for the reference's jax.grad, pinned to the numpy oracle in tests/test_autodiff_pins.py): the
  whole matrix dE/dh1 at 1e-8 relative (measured 1.2e-14) and the block energy at 1e-11, with one walker's
  energy clamped to E_est.
* The differentiation case (N = 128, n = 16, G = 640, 512 walkers, dt = 0.005, 200 steps, checkpoint every
  20; seed 3, every L_g / sqrt(N), D symmetric N(0,1) 0.05): <dE/dh1, D> against the tangent propagation
  (1e-8) and against central differences at lambda = +-1e-4 with common fields on the product's block path
  (1e-4), FP64 and FP32 contractions.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.cases import rel

pytestmark = pytest.mark.gpu


def _rand(rng, shape, cplx):
    x = rng.standard_normal(shape)
    return x + 1j * rng.standard_normal(shape) if cplx else x


def _op(x, op):
    return {"N": x, "T": np.swapaxes(x, -1, -2), "C": np.conj(np.swapaxes(x, -1, -2)), "R": np.conj(x)}[op]


@pytest.mark.parametrize("ac", [False, True])
@pytest.mark.parametrize("bc", [False, True])
def test_gemm_batched_all_ops(gpu, ac, bc):
    from trot_amd import _lib

    rng = np.random.default_rng(7)
    ops_a = ["N", "T", "C", "R"] if ac else ["N", "T"]
    ops_b = ["N", "T", "C", "R"] if bc else ["N", "T"]
    B, M, N, K = 3, 37, 70, 19
    for opa in ops_a:
        for opb in ops_b:
            a = _rand(rng, (B, M, K) if opa in "NR" else (B, K, M), ac)
            b = _rand(rng, (B, K, N) if opb in "NR" else (B, N, K), bc)
            ref = _op(a, opa) @ _op(b, opb)
            got = _lib.gemm_batched(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), opa, opb).cpu().numpy()
            assert rel(got, ref) < 1e-13, (opa, opb)
    # shared A (2-D), complex alpha, beta, and a strided output (a column block of a wider matrix)
    a = _rand(rng, (M, K), ac)
    b = _rand(rng, (B, K, N), True)
    wide = torch.from_numpy(_rand(rng, (B, M, 3 * N), True)).cuda()
    before = wide.cpu().numpy()
    _lib.gemm_batched(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), alpha=0.5 - 2j, beta=-1.5,
                      out=wide[:, :, N:2 * N])
    after = wide.cpu().numpy()
    assert rel(after[:, :, N:2 * N], (0.5 - 2j) * (a @ b) - 1.5 * before[:, :, N:2 * N]) < 1e-13
    assert np.array_equal(after[:, :, :N], before[:, :, :N]) and np.array_equal(after[:, :, 2 * N:], before[:, :, 2 * N:])


def test_gemm_batched_many_small_products(gpu):
    """70,000 batch entries: more than a 3D grid's 65,535, and not a multiple of the 8 XCDs the tiles are
    dealt to (the last group of eight workgroups is partly empty)."""
    from trot_amd import _lib

    rng = np.random.default_rng(9)
    B, M, N, K = 70000, 5, 3, 4
    a = _rand(rng, (B, M, K), True)
    b = _rand(rng, (B, K, N), False)
    got = _lib.gemm_batched(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), alpha=2.0).cpu().numpy()
    assert rel(got, 2.0 * (a @ b)) < 1e-13


def test_gemm_batched_refusals_and_inverse(gpu):
    from trot_amd import _lib

    x = torch.zeros((2, 3, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.TrotError):
        _lib.gemm_batched(x, x, opa="C")  # conjugate of a real operand
    with pytest.raises(ValueError):
        _lib.gemm_batched(x, torch.zeros((2, 4, 3), dtype=torch.float64, device="cuda"))
    rng = np.random.default_rng(8)
    m = _rand(rng, (9, 17, 17), True)
    minv, det = _lib.inverse_det(torch.from_numpy(m).cuda())
    assert rel(minv.cpu().numpy(), np.linalg.inv(m)) < 1e-12
    assert rel(det.cpu().numpy(), np.linalg.det(m)) < 1e-12


# ----------------------------------------------------------------------------------------- small system


def _small(seed=21, N=24, n=4, G=30, W=16):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((N, N)) / np.sqrt(N)
    h1 = np.triu(a) + np.triu(a, 1).T
    b = rng.standard_normal((G, N, N))
    chol = 0.12 * (b + np.swapaxes(b, 1, 2)) * np.exp(-np.arange(G) / G)[:, None, None]
    c = np.linalg.eigh(h1)[1][:, :n]
    w0 = c[None] + 0.05 * (rng.standard_normal((W, N, n)) + 1j * rng.standard_normal((W, N, n)))
    fields = [rng.standard_normal((W, G)) for _ in range(40)]
    Ds = []
    for _ in range(3):
        d = rng.standard_normal((N, N))
        Ds.append(0.05 * (d + d.T))
    return dict(h0=0.4, h1=h1, chol=chol, c=c, w0=w0, fields=fields, Ds=Ds, N=N, n=n, G=G, W=W)


def _adjoint_run(s, h1, steps, every=4, precision="f64", dt=0.01):
    from trot_amd import adjoint, prop

    params = prop.QmcParams(dt=dt, n_walkers=s["W"], seed=1, n_prop_steps=steps)
    ac = adjoint.build_ctx(s["h0"], h1, s["chol"], s["c"], params, precision=precision)
    f = [torch.from_numpy(x).cuda() for x in s["fields"]]
    return adjoint.run(ac, torch.from_numpy(s["w0"]).cuda(), lambda k: f[k], steps, checkpoint_every=every)


def _energy_only(s, h1, steps, precision="f64", dt=0.01):
    from trot_amd import adjoint, prop

    params = prop.QmcParams(dt=dt, n_walkers=s["W"], seed=1, n_prop_steps=steps)
    ac = adjoint.build_ctx(s["h0"], h1, s["chol"], s["c"], params, precision=precision)
    snap = adjoint.init_snap(ac, torch.from_numpy(s["w0"]).cuda())
    for k in range(steps):
        snap = adjoint.step(ac, snap, torch.from_numpy(s["fields"][k]).cuda())
    return adjoint._block_energy_terms(ac, snap)[0]


def test_adjoint_small_system(gpu):
    from trot_amd import adjoint, blocks, diff, meas, prop, trial
    from trot_amd.core import HamChol, System

    s = _small()
    steps = 12
    res = _adjoint_run(s, s["h1"], steps, every=4)
    grad = res.grad_h1.numpy()
    assert np.allclose(grad, grad.T)

    # the product's block path at the same h1 and fields
    N, n, W = s["N"], s["n"], s["W"]
    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(s["h0"], s["h1"], s["chol"]).to_device()
    tr = trial.RhfTrial(s["c"])
    mops, tops = meas.make_rhf_meas_ops(sys_), trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted")
    params = prop.QmcParams(dt=0.01, n_walkers=W, seed=1, n_prop_steps=steps)
    pctx, mctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params), mops.build_meas_ctx(ham, tr)
    st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=torch.from_numpy(s["w0"]).cuda())
    _, obs = blocks.block(st, sys=sys_, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                          meas_ctx=mctx, prop_ops=pops, prop_ctx=pctx,
                          fields_seq=[torch.from_numpy(x).cuda() for x in s["fields"][:steps]])
    assert abs(res.energy - float(obs.scalars["energy"].item())) <= 1e-11 * abs(res.energy)

    h = 1e-5
    f = [torch.from_numpy(x).cuda() for x in s["fields"]]
    for D in s["Ds"]:
        fd = (_energy_only(s, s["h1"] + h * D, steps) - _energy_only(s, s["h1"] - h * D, steps)) / (2 * h)
        ad = float(np.sum(grad * D))
        tc = diff.build_ctx(s["h0"], s["h1"], s["chol"], s["c"], D, 0.0, params)
        tg = diff.run(tc, torch.from_numpy(s["w0"]).cuda(), lambda k: f[k], steps).denergy
        print(f"[adjoint] <grad, D> {ad:.12e}  tangent {tg:.12e}  central difference {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-6 * abs(fd)
        assert abs(ad - tg) <= 1e-8 * abs(tg)
    # single entries: the (i, j) + (j, i) perturbation
    for (i, j) in [(0, 0), (3, 17), (23, 5)]:
        D = np.zeros((s["N"], s["N"]))
        D[i, j] += 1.0
        D[j, i] += 1.0 if i != j else 0.0
        fd = (_energy_only(s, s["h1"] + h * D, steps) - _energy_only(s, s["h1"] - h * D, steps)) / (2 * h)
        assert abs(float(np.sum(grad * D)) - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(grad).max()), (i, j)

    # checkpoint spacing: recomputation is deterministic
    for every in (1, 5, 100):
        other = _adjoint_run(s, s["h1"], steps, every=every)
        assert other.energy == res.energy and torch.equal(other.grad_h1, res.grad_h1), every

    # FP32 Cholesky contractions
    r32 = _adjoint_run(s, s["h1"], steps, every=4, precision="f32")
    assert abs(r32.energy - res.energy) <= 1e-5 * abs(res.energy)
    assert rel(r32.grad_h1.numpy(), grad) < 1e-4

    # through autograd: h1.grad is the response density, lam.grad = <grad, D>
    h1t = torch.tensor(s["h1"], dtype=torch.float64, requires_grad=True)
    e = adjoint.ResponseEnergy.apply(h1t, lambda h: res)
    e.backward()
    assert torch.equal(h1t.grad, res.grad_h1)
    lam = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    D0 = torch.from_numpy(s["Ds"][0])
    e = adjoint.ResponseEnergy.apply(torch.from_numpy(s["h1"]) + lam * D0, lambda h: res)
    e.backward()
    assert abs(lam.grad.item() - float(np.sum(grad * s["Ds"][0]))) <= 1e-14 * abs(lam.grad.item()) + 1e-300


def test_adjoint_against_autograd_of_the_restatement(gpu):
    from oracle import autodiff

    s = _small()
    steps = 12
    res = _adjoint_run(s, s["h1"], steps, every=4)
    h1t = torch.tensor(s["h1"], dtype=torch.float64, requires_grad=True)
    e, info = autodiff.block_energy(s["h0"], h1t, s["chol"], s["c"], s["w0"], s["fields"][:steps], 0.01)
    e.backward()
    g = h1t.grad.numpy()
    g = 0.5 * (g + g.T)
    err = rel(res.grad_h1.numpy(), g)
    print(f"[adjoint] vs autograd of the restatement: energy {res.energy:.14e} / {e.item():.14e}, "
          f"dE/dh1 rel {err:.2e}, masks {info}", flush=True)
    # masked importances and capped weights carry no derivative in either; a clamped energy (one walker here)
    # is differentiated through E_est = Re E_L(walker 0) by both
    assert info["masked_importance"] == 0 and info["capped_weight"] == 0
    assert abs(res.energy - e.item()) <= 1e-11 * abs(e.item())
    assert err < 1e-8  # measured 1.2e-14


# ---------------------------------------------------------------------------------- differentiation case

N4, n4, G4, W4, DT4, STEPS4, H4 = 128, 16, 640, 512, 0.005, 200, 1e-4


def _fields4(k):
    from trot_amd import synthetic

    return torch.from_numpy(synthetic.fields(3, k, W4, G4)).cuda()


def test_adjoint_differentiation_case(gpu):
    import time

    from trot_amd import adjoint, diff, prop, synthetic

    syn = synthetic.make_synthetic(N4, n4, G4, W4, 3, 256.0, chol_scale=1.0 / np.sqrt(N4))
    rng = np.random.default_rng(33)
    a = rng.standard_normal((N4, N4))
    D = 0.05 * (np.triu(a) + np.triu(a, 1).T)
    params = prop.QmcParams(dt=DT4, n_walkers=W4, seed=3, n_prop_steps=STEPS4)
    w0 = torch.from_numpy(syn.walkers).cuda()
    fcache = {}

    def fields(k):
        if k not in fcache:
            fcache[k] = _fields4(k)
        return fcache[k]

    out = {}
    for precision in ("f64", "f32"):
        ac = adjoint.build_ctx(syn.h0, syn.h1, syn.chol, syn.trial, params, precision=precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out[precision] = adjoint.run(ac, w0, fields, STEPS4, checkpoint_every=20)
        torch.cuda.synchronize()
        print(f"[adjoint] {precision}: 200 steps forward + reverse in {time.perf_counter() - t0:.2f} s", flush=True)
    ad = float(np.sum(out["f64"].grad_h1.numpy() * D))
    ad32 = float(np.sum(out["f32"].grad_h1.numpy() * D))
    tc = diff.build_ctx(syn.h0, syn.h1, syn.chol, syn.trial, D, 0.0, params)
    tg = diff.run(tc, w0, fields, STEPS4).denergy
    from tests.test_gpu_diff import _product_block_energy

    fd = (_product_block_energy(syn, D, H4) - _product_block_energy(syn, D, -H4)) / (2 * H4)
    print(f"[adjoint] E {out['f64'].energy:.10f}  <grad, D> f64 {ad:.10e} f32 {ad32:.10e}  tangent {tg:.10e}  "
          f"central difference {fd:.10e}", flush=True)
    assert abs(ad - tg) <= 1e-8 * abs(tg)
    assert abs(ad - fd) <= 1e-4 * abs(fd)
    assert abs(ad32 - fd) <= 1e-4 * abs(fd)
    assert rel(out["f32"].grad_h1.numpy(), out["f64"].grad_h1.numpy()) < 1e-4
