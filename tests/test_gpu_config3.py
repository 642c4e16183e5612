"""The precision form's workload at its full size: N = 512, n = 64, G = 2560, 2048 walkers, dt = 0.005 (seed 2,
tau 512), every L_gamma scaled by 1/sqrt(N) so that the walkers keep O(1) weights (the recipe as written
overflows within a few steps, in the oracle as on the GPU).

1. FP64: all 2048 walkers take 3 steps on the GPU.  A sample of 64 walkers (walker 0, whose energy sets the
   initial e_estimate, and 63 others) takes the same 3 steps in the CPU oracle with identical fields; the
   walkers couple only through the population-control shift, which the oracle takes from the GPU run before
   each step.  Walkers, weights and overlaps agree at 1e-10 (weights and overlaps elementwise).
2. The precision forms against the FP64 GPU path (pinned by 1.) over 20 paired steps with identical fields:
   split-BF16 with Taylor terms 3-6 as one bf16 pass (the benchmark's form), split-BF16 with all three passes,
   FP32, and the reference's mixed mode (mixed_precision=True: FP32 VHS, FP64 Taylor).  Each form's local
   energies come from its own exchange kernels on its own walkers: the block energy is within 0.1 mHa of FP64,
   every walker's E_L within 0.05 mHa, every weight within 1e-4 relative.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import afqmc_oracle as O
from tests.cases import rel, rel_each

pytestmark = pytest.mark.gpu

WORKLOAD = "n512_o64_g2560_w2048"


@pytest.fixture(scope="module")
def big(gpu):
    from trot_amd import synthetic, trial
    from trot_amd.core import HamChol, System
    from trot_amd import prop

    cfg = synthetic.WORKLOADS[WORKLOAD]
    N, n, G, W = cfg["norb"], cfg["nocc"], cfg["nchol"], cfg["n_walkers"]
    print(f"[config3] building the synthetic system N={N} n={n} G={G} W={W}", flush=True)
    syn = synthetic.make_synthetic(N, n, G, W, cfg["seed"], cfg["tau"], device=gpu, chol_scale=1.0 / np.sqrt(N))
    sys_ = System(N, (n, n), "restricted")
    env = dict(cfg=cfg, syn=syn, sys=sys_, ham=HamChol(syn.h0, torch.from_numpy(syn.h1).to(gpu), syn.chol),
               tr=trial.RhfTrial(syn.trial), tops=trial.make_rhf_trial_ops(sys_),
               params=prop.QmcParams(dt=cfg["dt"], n_walkers=W, seed=cfg["seed"]))
    yield env
    del env, syn
    torch.cuda.empty_cache()


class _Form:
    def __init__(self, env, precision=None, taylor_precision=None, hi_from=0, mixed_precision=False):
        from trot_amd import meas, prop

        self.env = env
        meas_prec = "f64" if mixed_precision else (precision or "f64")
        self.mops = meas.make_rhf_meas_ops(env["sys"], precision=meas_prec)
        self.pops = prop.make_prop_ops("restricted", "restricted", mixed_precision, precision=precision,
                                       taylor_precision=taylor_precision, taylor_hi_from=hi_from)
        self.pctx = self.pops.build_prop_ctx(env["ham"], env["tops"].get_rdm1(env["tr"]), env["params"])
        self.mctx = self.mops.build_meas_ctx(env["ham"], env["tr"])

    def init(self):
        e = self.env
        return self.pops.init_prop_state(sys=e["sys"], ham_data=e["ham"], trial_ops=e["tops"], trial_data=e["tr"],
                                         meas_ops=self.mops, params=e["params"], initial_walkers=e["syn"].walkers)

    def step(self, st, f):
        e = self.env
        return self.pops.step(st, params=e["params"], ham_data=e["ham"], trial_data=e["tr"], trial_ops=e["tops"],
                              meas_ops=self.mops, meas_ctx=self.mctx, prop_ctx=self.pctx, fields=f)

    def energies(self, st):
        from trot_amd.meas import _h0

        _, g = self.mops.engine.overlap_green(st.walkers, self.env["tr"])
        return self.mops.engine.energy_from_green(g, _h0(self.env["ham"]), self.mctx).real


def test_config3_fp64_against_the_oracle(big):
    from trot_amd import synthetic

    cfg, syn = big["cfg"], big["syn"]
    N, G, W, dt = cfg["norb"], cfg["nchol"], cfg["n_walkers"], cfg["dt"]
    gpu = _Form(big)
    st = gpu.init()
    idx = np.concatenate([[0], np.sort(np.random.default_rng(5).choice(np.arange(1, W), 63, replace=False))])
    print("[config3] oracle contexts (N=512, G=2560)", flush=True)
    chol = syn.chol.cpu().numpy()
    ref_ops = O.make_rhf_ops(syn.h0, syn.h1, chol, syn.trial)
    ref_ctx = O.build_prop_ctx(syn.h0, syn.h1, chol, O.rhf_rdm1(syn.trial), dt)
    del chol
    w0 = syn.walkers.cpu().numpy()[idx]
    ref = O.init_prop_state(w0, ref_ops, len(idx))
    assert abs(ref.e_estimate - st.e_estimate.item()) <= 1e-10 * abs(ref.e_estimate)
    assert rel_each(st.overlaps.cpu().numpy()[idx], ref.overlaps) < 1e-10
    params = O.QmcParams(dt=dt, n_walkers=len(idx))
    for s in range(3):
        f = synthetic.fields(cfg["seed"], s, W, G)
        ref.pop_control_ene_shift = float(st.pop_control_ene_shift.item())  # the all-walker shift
        st = gpu.step(st, torch.from_numpy(f).cuda())
        ref = O.afqmc_step(ref, f[idx], ref_ctx, ref_ops, params)
        print(f"[config3] step {s + 1}: oracle sample done", flush=True)
        wg = st.walkers.cpu().numpy()[idx]
        assert rel(wg, ref.walkers) < 1e-10, s
        assert rel_each(st.weights.cpu().numpy()[idx], ref.weights) < 1e-10, s
        assert rel_each(st.overlaps.cpu().numpy()[idx], ref.overlaps) < 1e-10, s
        assert float(st.weights.min()) > 0.5 and float(st.weights.max()) < 2.0  # the scaled system stays O(1)


def test_config3_precision_forms_against_fp64(big):
    cfg = big["cfg"]
    G, W = cfg["nchol"], cfg["n_walkers"]
    forms = {
        "bf16x3, Taylor 3-6 one pass": _Form(big, "bf16x3", "bf16x3", hi_from=3),
        "bf16x3, all three passes": _Form(big, "bf16x3", "bf16x3"),
        "f32": _Form(big, "f32", "f32"),
        "reference mixed (mixed_precision=True)": _Form(big, mixed_precision=True),
    }
    assert forms["reference mixed (mixed_precision=True)"].pctx.taylor_precision == "f64"
    ref = _Form(big)
    states = {k: v.init() for k, v in forms.items()}
    st64 = ref.init()
    for s in range(20):
        g = torch.Generator(device="cuda")
        g.manual_seed(7000 + s)
        f = torch.randn((W, G), dtype=torch.float64, device="cuda", generator=g)
        st64 = ref.step(st64, f)
        for k in forms:
            states[k] = forms[k].step(states[k], f)
    e64 = ref.energies(st64)
    blk64 = float((st64.weights * e64).sum() / st64.weights.sum())
    for k, form in forms.items():
        st = states[k]
        e = form.energies(st)
        blk = float((st.weights * e).sum() / st.weights.sum())
        d_walker = 1e3 * float((e - e64).abs().max())
        d_block = 1e3 * abs(blk - blk64)
        d_weight = float(((st.weights - st64.weights).abs() / st64.weights.abs()).max())
        print(f"[config3] {k}: per-walker max |dE| {d_walker:.5f} mHa, block dE {d_block:.6f} mHa, "
              f"max rel weight diff {d_weight:.2e}", flush=True)
        assert d_block <= 0.1, k
        assert d_walker <= 0.05, k
        assert d_weight <= 1e-4, k
