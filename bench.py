"""Throughput of the phaseless AFQMC walker step on one MI355X: mixed-precision MFMA vs FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--precision bf16x3|f32] [--taylor-precision f32|bf16x3|f64] [--no-cpu] [--no-block]

Workload (default): the nbasis=512 synthetic system, N=512 orbitals, n=64
occupied, G=2560 Cholesky vectors, 2048 complex128 walkers, dt=0.005, RHF
trial, restricted walkers (seeded synthetic Hamiltonian, trot_amd/synthetic.py,
every L_gamma scaled by 1/sqrt(N) so that the walkers keep O(1) weights: with
the recipe's unscaled L every importance leaves [floor, cap] within ~10 steps
and the overlaps overflow, in the CPU oracle as on the GPU; the rate on the
unscaled recipe is reported beside the headline).  A step is one phaseless step
for all walkers: field draw, force bias, shifted fields, VHS build, one-body /
Taylor / one-body propagation, overlaps and Green's functions, weight update.
All inputs are resident in HBM before the timed regions.

The same process times K steps of the FP64 path and K steps of each reduced-
precision form (split-BF16 with Taylor terms 3-6 as one bf16 pass: `value`;
split-BF16 with every product three-pass; FP32; the reference's own mixed mode,
FP32 VHS with FP64 Taylor).  Each form's accuracy is measured against the FP64
path (which the test suite pins to the CPU oracle) by running all of them in
lockstep from one state with identical fields and comparing the per-walker
local energies and the weighted block energy, in mHa.

`roofline` describes the dominant kernel of the mixed step (the VHS GEMM);
its duration comes from HIP events recorded around its launches inside the
timed steps.  `cpu_baseline` is the NumPy FP64 oracle timed on a bounded
sample of the same workload on the host cores.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0}  # dense MFMA TFLOP/s, MI355X datasheet
METRIC = "walker-steps/s at nbasis=512: MFMA mixed vs FP64 step"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="n512_o64_g2560_w2048")
    ap.add_argument("--precision", default="bf16x3", choices=["f32", "bf16x3"])
    ap.add_argument("--taylor-precision", default="bf16x3", choices=["f64", "f32", "bf16x3"])
    ap.add_argument("--taylor-hi-from", type=int, default=3,
                    help="first Taylor term computed as one bf16 pass in the split-BF16 propagation (0: none)")
    ap.add_argument("--accuracy-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-block", action="store_true", help="skip the block-tail timing")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 leg (profiling runs)")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-replayed step (profiling runs)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the other precision forms (all-pass split BF16, FP32, the reference's mixed mode)")
    ap.add_argument("--no-unscaled", action="store_true", help="skip the rate on the unscaled recipe")
    ap.add_argument("--cpu-walkers", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--cisd-workload", default="cisd_n128_o16_g512_w1024")
    ap.add_argument("--cisd-only", action="store_true",
                    help="only the CISD-trial rates, printed as their own JSON line (profiling runs)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batching form's rates (32 fragments; profiling runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device copies of 1 GiB (scale reference for rocprofv3 FETCH/WRITE_SIZE)")
    return ap.parse_args()


def calibrate_copy() -> None:
    """Three copies of a 1 GiB float64 tensor: 1 GiB read + 1 GiB written per copy kernel."""
    import torch

    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8, "copies": 3}))


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(tag: str):
    """HBM bytes per launch of the dominant kernel from a committed rocprofv3 --pmc summary, if any."""
    p = ROOT / "profiles" / f"{tag}_traffic.md"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


class Runner:
    """One precision configuration of the step on a shared synthetic system."""

    def __init__(self, env, precision: str | None, taylor_precision: str | None, hi_from: int = 0,
                 mixed_precision: bool = False):
        from trot_amd import meas, prop

        self.env = env
        # the reference's mixed mode keeps the RHF force bias and energy in FP64 (trot/meas/rhf.py has no
        # precision switch); the explicit forms run them in the VHS's precision
        self.meas_precision = "f64" if mixed_precision else precision
        self.mops = meas.make_rhf_meas_ops(env["sys"], precision=self.meas_precision)
        self.pops = prop.make_prop_ops("restricted", "restricted", mixed_precision, precision=precision,
                                       taylor_precision=taylor_precision, taylor_hi_from=hi_from)
        self.pctx = self.pops.build_prop_ctx(env["ham"], env["tops"].get_rdm1(env["tr"]), env["params"])
        self.mctx = self.mops.build_meas_ctx(env["ham"], env["tr"])
        self.precision = self.pctx.precision
        self.taylor_precision = self.pctx.taylor_precision

    def init(self):
        e = self.env
        return self.pops.init_prop_state(sys=e["sys"], ham_data=e["ham"], trial_ops=e["tops"], trial_data=e["tr"],
                                         meas_ops=self.mops, params=e["params"], initial_walkers=e["w0"])

    def step(self, st, fields=None):
        e = self.env
        return self.pops.step(st, params=e["params"], ham_data=e["ham"], trial_data=e["tr"], trial_ops=e["tops"],
                              meas_ops=self.mops, meas_ctx=self.mctx, prop_ctx=self.pctx, fields=fields)

    def timed_graph(self, st, warmup: int, steps: int, tag: str):
        """Steps replayed from one recorded HIP graph (trot_amd.graph.GraphStep)."""
        import torch

        from trot_amd.graph import GraphStep

        e = self.env
        gs = GraphStep(self.pops)
        kw = dict(params=e["params"], ham_data=e["ham"], trial_data=e["tr"], trial_ops=e["tops"],
                  meas_ops=self.mops, meas_ctx=self.mctx, prop_ctx=self.pctx)
        st = gs(st, **kw)  # eager warm-up step + recording
        for i in range(warmup):
            st = gs(st, **kw)
        torch.cuda.synchronize()
        log(f"{tag}: recorded, {warmup} warm-up replays done")
        t0 = time.perf_counter()
        for _ in range(steps):
            st = gs(st, **kw)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log(f"{tag}: {steps} steps in {dt:.3f}s ({1e3 * dt / steps:.2f} ms/step)")
        del gs
        return st, dt

    def timed(self, st, warmup: int, steps: int, tag: str):
        import torch

        from trot_amd import prop

        for i in range(warmup):
            st = self.step(st)
            torch.cuda.synchronize()
            log(f"{tag}: warmup step {i + 1}/{warmup} done")
        timer = prop.KernelTimer()
        self.pctx.low["timer"] = timer
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = self.step(st)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        self.pctx.low.pop("timer")
        log(f"{tag}: {steps} steps in {dt:.3f}s ({1e3 * dt / steps:.2f} ms/step)")
        return st, dt, timer.mean_ms("vhs")


FORMS = (  # (name, precision, taylor_precision, taylor_hi_from, mixed_precision flag)
    ("bf16x3_all_passes", "bf16x3", "bf16x3", 0, False),
    ("f32", "f32", "f32", 0, False),
    ("reference_mixed", None, None, 0, True),
)


def main():
    args = parse()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU only: trot_amd has no multi-GPU path (--gpus must be 1)")
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.cisd_only:
        print(json.dumps(cisd_rates(args)), flush=True)
        return
    if args.batch_only:
        print(json.dumps(batch_rates()), flush=True)
        return
    import torch

    from trot_amd import _lib, meas, prop, synthetic, trial
    from trot_amd.core import HamChol, System

    _lib.lib()
    cfg = synthetic.WORKLOADS[args.workload]
    N, n, G, W, dt = cfg["norb"], cfg["nocc"], cfg["nchol"], cfg["n_walkers"], cfg["dt"]
    dev = torch.device("cuda")
    t0 = time.perf_counter()
    syn = synthetic.make_synthetic(N, n, G, W, cfg["seed"], cfg["tau"], device=dev, chol_scale=1.0 / np.sqrt(N))
    log(f"synthetic system N={N} n={n} G={G} W={W} (L / sqrt(N)) in {time.perf_counter() - t0:.1f}s")
    sys_ = System(N, (n, n), "restricted")
    env = dict(sys=sys_, ham=HamChol(syn.h0, torch.from_numpy(syn.h1).to(dev), syn.chol),
               tr=trial.RhfTrial(syn.trial), tops=trial.make_rhf_trial_ops(sys_),
               params=prop.QmcParams(dt=dt, n_walkers=W, seed=cfg["seed"]), w0=syn.walkers)

    mixed = Runner(env, args.precision, args.taylor_precision, args.taylor_hi_from)
    st_mix = mixed.init()
    torch.cuda.synchronize()
    log(f"mixed contexts ready at {time.perf_counter() - t0:.1f}s (sym-block VHS: {bool(mixed.pctx.low.get('symblk'))})")
    st_mix, t_mix, vhs_ms = mixed.timed(st_mix, args.warmup, args.steps, f"mixed {args.precision}")
    value = W * args.steps / t_mix
    w_min, w_max = float(st_mix.weights.min()), float(st_mix.weights.max())
    log(f"weights after the timed steps in [{w_min:.4f}, {w_max:.4f}]")
    # the same step replayed from one HIP graph per step (fields drawn on the device)
    rate_graph = None
    if not args.no_graph:
        st_mix, t_mix_graph = mixed.timed_graph(st_mix, args.warmup, args.steps, f"mixed {args.precision} graph")
        rate_graph = W * args.steps / t_mix_graph
    del st_mix

    fp64 = None
    rate64 = None
    if not args.no_fp64:
        fp64 = Runner(env, "f64", "f64")
        st64 = fp64.init()
        st64, t64, vhs64_ms = fp64.timed(st64, args.warmup, args.steps, "fp64")
        rate64 = W * args.steps / t64
        del st64

    # the other precision forms: rates, then all of them in lockstep with FP64 for the energy error
    forms = {}
    runners = {"headline": mixed}
    if not args.no_forms:
        for name, prec, tprec, hi, mp in FORMS:
            r = Runner(env, prec, tprec, hi, mixed_precision=mp)
            st = r.init()
            st, t_f, _ = r.timed(st, args.warmup, args.steps, name)
            del st
            rate = W * args.steps / t_f
            forms[name] = {"precision": r.pctx.precision, "taylor_precision": r.pctx.taylor_precision,
                           "meas_precision": r.meas_precision, "walker_steps_per_s": round(rate, 2),
                           "ms_per_step": round(1e3 * t_f / args.steps, 3),
                           "speedup_vs_fp64": None if rate64 is None else round(rate / rate64, 3)}
            runners[name] = r

    accuracy = None
    if fp64 is not None and args.accuracy_steps > 0:
        acc = paired_accuracy(env, runners, fp64, args.accuracy_steps)
        accuracy = acc.pop("headline")
        for name in forms:
            forms[name]["accuracy_vs_fp64"] = acc[name]
    for name in forms:
        del runners[name]
    del runners

    unscaled = None
    if not args.no_unscaled:
        # the recipe as written (no 1/sqrt(N)): same flops, walkers overflow within a few steps
        env_u = dict(env, ham=HamChol(syn.h0, env["ham"].h1, syn.chol * float(np.sqrt(N))))
        r_u = Runner(env_u, args.precision, args.taylor_precision, args.taylor_hi_from)
        st_u = r_u.init()
        st_u, t_u, _ = r_u.timed(st_u, args.warmup, args.steps, "mixed, unscaled recipe")
        unscaled = {"system": "the recipe as written (L_gamma without 1/sqrt(N))",
                    "mixed_ms_per_step": round(1e3 * t_u / args.steps, 3),
                    "mixed_walker_steps_per_s": round(W * args.steps / t_u, 2),
                    "finite_weights_after": int(torch.isfinite(st_u.weights).sum()),
                    "weights_in_floor_cap_after": int(((st_u.weights > 0) & (st_u.weights < 100)).sum())}
        del r_u, st_u, env_u
        torch.cuda.empty_cache()

    # roofline of the dominant kernel (VHS GEMM of the mixed step): flops the algorithm run here needs
    sym = bool(mixed.pctx.low.get("symblk"))
    cols = _lib.symblk_cols(N) if sym else N * N
    passes = 3 if args.precision == "bf16x3" else 1
    vhs_flop = 4.0 * W * G * cols  # real x complex MAC = 4 flop over the stored V entries
    achieved = vhs_flop * passes / (vhs_ms * 1e-3) / 1e12
    peak = PEAK[args.precision]
    tag = f"vhs_{args.precision}_{args.workload}"
    roofline = {"bound": "mfma", "kernel": f"VHS GEMM ({args.precision}, {'sym-block' if sym else 'dense'} V)",
                "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                "traffic": traffic_from_profiles(tag),
                "traffic_source": f"profiles/{tag}_traffic.md: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over "
                                  "this bench command, each scaled by the counter's reading for a 1 GiB device copy "
                                  "(not measured inside this run)",
                "launch_ms": round(vhs_ms, 3), "flop_per_launch": vhs_flop * passes}

    # the whole step against its ideal: T* = sum_k F_k / peak_k over the contractions as run here
    # (dense complex products 8 flop / MAC, real x complex 4; split-BF16 passes counted against the BF16
    # peak, FP64 ones against FP64), then T* / measured
    ps = PEAK[args.precision] * 1e12
    p64 = PEAK["f64"] * 1e12
    tp = PEAK[args.taylor_precision] * 1e12
    hi = args.taylor_hi_from if args.taylor_precision == "bf16x3" and args.taylor_hi_from > 0 else 0
    k_terms = env["params"].n_exp_terms
    t_passes = (3 * (min(hi, k_terms + 1) - 1) + (k_terms - hi + 1)) if hi else (3 if args.taylor_precision ==
                                                                                 "bf16x3" else 1) * k_terms
    ob_passes = 3 if args.taylor_precision == "bf16x3" else 1
    terms = {
        "vhs": vhs_flop * passes / ps,
        "force_bias": 4.0 * W * G * N * n * passes / ps,
        "taylor": 8.0 * W * N * N * n * t_passes / tp,
        "one_body": 2 * 4.0 * W * N * N * n * ob_passes / tp,
        "overlap_inverse_green_f64": (4.0 * W * N * n * n + 8.0 * W * N * n * n + 8.0 * W * n ** 3) / p64,
    }
    t_star = sum(terms.values())
    step_roofline = {"t_star_ms": round(1e3 * t_star, 3), "ms_per_step": round(1e3 * t_mix / args.steps, 3),
                     "frac": round(t_star / (t_mix / args.steps), 4),
                     "terms_ms": {k: round(1e3 * v, 3) for k, v in terms.items()},
                     "taylor_passes": t_passes}
    if rate64 is not None:
        f64_flop = (4.0 * W * G * cols + 4.0 * W * G * N * n + 8.0 * W * N * N * n * k_terms
                    + 2 * 4.0 * W * N * N * n + 4.0 * W * N * n * n + 8.0 * W * N * n * n + 8.0 * W * n ** 3)
        step_roofline["fp64_t_star_ms"] = round(1e3 * f64_flop / p64, 3)
        step_roofline["fp64_frac"] = round(f64_flop / p64 / (W / rate64), 4)

    block_rate = None
    block_graph = None
    small = None
    fp_rates = None
    if not args.no_block:
        st_blk = mixed.init()
        for _ in range(2):
            st_blk = mixed.step(st_blk)
        block_rate = block_inclusive(env, mixed, st_blk, t_mix / args.steps)
        block_graph = graph_block_rate(env, mixed, st_blk)
        del st_blk
        small = launch_bound_case()
        fp_rates = free_projection_rates(syn, env, args)

    cpu = None if args.no_cpu else cpu_baseline(syn, mixed, args, dt)

    cisd = None if args.no_block else cisd_rates(args)
    differentiation = None if args.no_block else differentiation_rates()
    batched = None if args.no_block else batch_rates()

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mix / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"mixed: {args.precision} MFMA Cholesky GEMMs (VHS, force bias, exchange), {args.taylor_precision} "
                 f"one-body products and Taylor terms"
                 + (f" 1-{args.taylor_hi_from - 1}, one-pass bf16 Taylor terms {args.taylor_hi_from}-6"
                    if args.taylor_hi_from > 0 else "")
                 + "; f64 walkers, det, inverse, QR, weights and energy sums",
        "data": "synthetic (seeded numpy PCG64 Hamiltonian, trial and walkers, every L_gamma / sqrt(N); "
                "on-device N(0,1) fields)",
        "config": {"workload": args.workload, "nbasis": N, "nocc": n, "nchol": G, "n_walkers": W, "dt": dt,
                   "n_exp_terms": env["params"].n_exp_terms, "trial": "RHF", "walker_kind": "restricted",
                   "precision": args.precision, "taylor_precision": args.taylor_precision,
                   "taylor_hi_from": args.taylor_hi_from, "chol_scale": "1/sqrt(N)",
                   "vhs_layout": "sym-block" if sym else "dense"},
        "weights_after_timed_steps": [round(w_min, 6), round(w_max, 6)],
        "fp64_walker_steps_per_s": None if rate64 is None else round(rate64, 2),
        "fp64_ms_per_step": None if rate64 is None else round(1e3 * W / rate64, 3),
        "speedup_vs_fp64": None if rate64 is None else round(value / rate64, 3),
        "accuracy_vs_fp64": accuracy,
        "precision_forms": forms,
        "unscaled_recipe": unscaled,
        "graph_step_walker_steps_per_s": None if rate_graph is None else round(rate_graph, 2),
        "block_inclusive_walker_steps_per_s": None if block_rate is None else round(block_rate, 2),
        "block_graph": block_graph,
        "launch_bound_case": small,
        "free_projection": fp_rates,
        "cisd": cisd,
        "differentiation": differentiation,
        "batch": batched,
        "roofline": roofline,
        "step_roofline": step_roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def paired_accuracy(env, forms: dict, fp64: Runner, steps: int, every: int = 50) -> dict:
    """Every form and the FP64 path from one initial state with identical fields, compared every ``every``
    steps.

    Each form's local energies come from its own (reduced-precision) exchange kernels on its own walkers;
    the reference values are the FP64 energies of the FP64 walkers.  The block energy is the weighted mean
    of Re E_L with each path's own weights.  Reported per form: the largest per-walker and block-energy
    differences over all checkpoints, and the values at the last one.
    """
    import torch

    from trot_amd.meas import _h0

    W = env["params"].n_walkers
    G = fp64.pctx.chol_flat.shape[0]
    h0 = _h0(env["ham"])
    states = {k: r.init() for k, r in forms.items()}
    b = fp64.init()
    worst = {k: {"walker": 0.0, "block": 0.0} for k in forms}
    last = {k: {} for k in forms}

    def energies(r, st):
        _, g = r.mops.engine.overlap_green(st.walkers, env["tr"])
        return r.mops.engine.energy_from_green(g, h0, r.mctx).real

    for s in range(steps):
        g = torch.Generator(device="cuda")
        g.manual_seed(9000 + s)
        f = torch.randn((W, G), dtype=torch.float64, device="cuda", generator=g)
        b = fp64.step(b, f)
        for k, r in forms.items():
            states[k] = r.step(states[k], f)
        if (s + 1) % every and s + 1 != steps:
            continue
        eb = energies(fp64, b)
        blk_b = float((b.weights * eb).sum() / b.weights.sum())
        for k, r in forms.items():
            a = states[k]
            ea = energies(r, a)
            blk_a = float((a.weights * ea).sum() / a.weights.sum())
            dw = 1e3 * float((ea - eb).abs().max())
            worst[k]["walker"] = max(worst[k]["walker"], dw)
            worst[k]["block"] = max(worst[k]["block"], 1e3 * abs(blk_a - blk_b))
            last[k] = {"per_walker_max_abs_dE_mHa": round(dw, 6),
                       "per_walker_rms_dE_mHa": round(1e3 * float(((ea - eb) ** 2).mean().sqrt()), 6),
                       "block_energy_dE_mHa": round(1e3 * abs(blk_a - blk_b), 6),
                       "block_energy_fp64_Ha": round(blk_b, 8),
                       "max_rel_weight_diff": float(((a.weights - b.weights).abs()
                                                     / b.weights.abs().clamp_min(1e-300)).max())}
            log(f"accuracy of {k} at paired step {s + 1}: {last[k]}")
    return {k: {"steps": steps, "checkpoint_every": every,
                "max_over_checkpoints": {"per_walker_max_abs_dE_mHa": round(worst[k]["walker"], 6),
                                         "block_energy_dE_mHa": round(worst[k]["block"], 6)},
                "final": last[k],
                "energies": "E_L from the form's own exchange kernels on its own walkers vs FP64 E_L of FP64 walkers"}
            for k in forms}


def block_inclusive(env, r: Runner, st, t_step: float) -> float:
    """Rate with the block tail (QR, energy, clamp / mean / EMA, SR) every n_prop_steps steps."""
    import torch

    from trot_amd import blocks

    params = env["params"]
    torch.cuda.synchronize()
    tb = time.perf_counter()
    # a block with zero propagation steps is exactly the tail: QR, overlaps, energy, SR
    import dataclasses

    p0 = dataclasses.replace(params, n_prop_steps=0)
    st, _ = blocks.block(st, sys=env["sys"], params=p0, ham_data=env["ham"], trial_data=env["tr"],
                         trial_ops=env["tops"], meas_ops=r.mops, meas_ctx=r.mctx, prop_ops=r.pops, prop_ctx=r.pctx)
    torch.cuda.synchronize()
    t_tail = time.perf_counter() - tb
    nps = params.n_prop_steps
    rate = params.n_walkers * nps / (nps * t_step + t_tail)
    log(f"block tail {1e3 * t_tail:.1f} ms; block-inclusive {rate:.1f} walker-steps/s")
    return rate


def graph_block_rate(env, r: Runner, st) -> dict:
    """One whole block (n_prop_steps steps + tail) replayed from one HIP graph (trot_amd.graph.GraphBlock)."""
    import torch

    from trot_amd.graph import GraphBlock

    params = env["params"]
    kw = dict(sys=env["sys"], params=params, ham_data=env["ham"], trial_data=env["tr"], trial_ops=env["tops"],
              meas_ops=r.mops, meas_ctx=r.mctx, prop_ops=r.pops, prop_ctx=r.pctx)
    gb = GraphBlock()
    t0 = time.perf_counter()
    st, _ = gb(st, **kw)
    torch.cuda.synchronize()
    t_rec = time.perf_counter() - t0
    t0 = time.perf_counter()
    st, obs = gb(st, **kw)
    torch.cuda.synchronize()
    t_blk = time.perf_counter() - t0
    rate = params.n_walkers * params.n_prop_steps / t_blk
    log(f"graph block: eager block + recording {t_rec:.2f}s, replay {1e3 * t_blk:.1f} ms "
        f"({rate:.1f} walker-steps/s incl. tail), energy {float(obs.scalars['energy']):.6f}")
    del gb
    return {"walker_steps_per_s": round(rate, 2), "ms_per_block": round(1e3 * t_blk, 2),
            "n_prop_steps": params.n_prop_steps}


def launch_bound_case() -> dict:
    """FP64 step and block at N=12, n=6, G=72, W=64 (the H12 STO-6G sizes): eager launches vs HIP graphs.

    The synthetic system of that size stands in for the chain (same shapes, same kernels)."""
    import torch

    from trot_amd import prop, synthetic, trial
    from trot_amd.core import HamChol, System
    from trot_amd.graph import GraphBlock, GraphStep

    N, n, G, W = 12, 6, 72, 64
    syn = synthetic.make_synthetic(N, n, G, W, 5, 64.0, device=torch.device("cuda"), chol_scale=1.0 / np.sqrt(N))
    sys_ = System(N, (n, n), "restricted")
    env = dict(sys=sys_, ham=HamChol(syn.h0, torch.from_numpy(syn.h1).cuda(), syn.chol),
               tr=trial.RhfTrial(syn.trial), tops=trial.make_rhf_trial_ops(sys_),
               params=prop.QmcParams(dt=0.005, n_walkers=W, seed=5, n_prop_steps=50), w0=syn.walkers)
    r = Runner(env, "f64", "f64")
    kw = dict(params=env["params"], ham_data=env["ham"], trial_data=env["tr"], trial_ops=env["tops"],
              meas_ops=r.mops, meas_ctx=r.mctx, prop_ctx=r.pctx)
    steps = 200
    st = r.init()
    for _ in range(5):
        st = r.step(st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = r.step(st)
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / steps
    gs = GraphStep(r.pops)
    for _ in range(5):
        st = gs(st, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = gs(st, **kw)
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / steps
    del gs
    from trot_amd.blocks import block

    bkw = dict(kw, sys=sys_, prop_ops=r.pops)
    st, _ = block(st, **bkw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, _ = block(st, **bkw)
    torch.cuda.synchronize()
    tb_eager = time.perf_counter() - t0
    gb = GraphBlock()
    st, _ = gb(st, **bkw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n_blk = 5
    for _ in range(n_blk):
        st, _ = gb(st, **bkw)
    torch.cuda.synchronize()
    tb_graph = (time.perf_counter() - t0) / n_blk
    out = {"workload": "n12_o6_g72_w64 FP64 (H12 STO-6G sizes, synthetic)", "eager_ms_per_step": round(1e3 * t_eager, 4),
           "graph_ms_per_step": round(1e3 * t_graph, 4), "eager_ms_per_block": round(1e3 * tb_eager, 3),
           "graph_ms_per_block": round(1e3 * tb_graph, 3), "n_prop_steps": 50,
           "graph_block_walker_steps_per_s": round(W * 50 / tb_graph, 1)}
    log(f"launch-bound case: {out}")
    return out


def free_projection_rates(syn, env, args) -> dict:
    """Free-projection step (trot_amd.fp, ten Taylor terms, QR norm and comb every step) at the workload's
    sizes, mixed and FP64, on the workload's (scaled) system."""
    import torch

    from trot_amd import fp, meas
    from trot_amd.core import HamChol

    N = env["sys"].norb
    ham = env["ham"]  # syn.chol already carries the 1/sqrt(N)
    W = env["params"].n_walkers
    params = fp.QmcParamsFp(dt=env["params"].dt, n_walkers=W, seed=11, ene0=0.0)
    out = {"workload": args.workload, "n_exp_terms": fp.FP_TAYLOR_TERMS}
    for tag, prec, tprec, hi in (("mixed", args.precision, args.taylor_precision, args.taylor_hi_from),
                                 ("fp64", "f64", "f64", 0)):
        mops = meas.make_rhf_meas_ops(env["sys"], precision=prec)
        pops = fp.make_prop_ops_fp("restricted", "restricted", env["sys"], precision=prec, taylor_precision=tprec,
                                   taylor_hi_from=hi)
        pctx = pops.build_prop_ctx(ham, env["tops"].get_rdm1(env["tr"]), params)
        mctx = mops.build_meas_ctx(ham, env["tr"])
        st = pops.init_prop_state(sys=env["sys"], ham_data=ham, trial_ops=env["tops"], trial_data=env["tr"],
                                  meas_ops=mops, params=params, initial_walkers=env["w0"])
        kw = dict(params=params, ham_data=ham, trial_data=env["tr"], trial_ops=env["tops"], meas_ops=mops,
                  meas_ctx=mctx, prop_ctx=pctx)
        st = pops.step(st, **kw)
        torch.cuda.synchronize()
        steps = 3
        t0 = time.perf_counter()
        for _ in range(steps):
            st = pops.step(st, **kw)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[f"{tag}_ms_per_step"] = round(1e3 * dt, 3)
        out[f"{tag}_walker_steps_per_s"] = round(W / dt, 1)
        log(f"free projection {tag}: {1e3 * dt:.2f} ms/step")
        del st, pctx, mctx, pops, mops
    return out


def _fragment(i, N, n, G, W, dt, chol_scale, precision="f64"):
    import torch

    from trot_amd import batch, meas, prop, synthetic, trial
    from trot_amd.core import HamChol, System

    syn = synthetic.make_synthetic(N, n, G, W, 100 + i, 256.0, nocc_b=n - 1, chol_scale=chol_scale)
    sys_ = System(N, (n, n - 1), "unrestricted")
    ham = HamChol(syn.h0, syn.h1, syn.chol).to_device()
    tr = trial.UhfTrial(syn.trial, syn.trial_b)
    mops, tops = meas.make_uhf_meas_ops(sys_, precision=precision), trial.make_uhf_trial_ops(sys_)
    if precision == "f64":
        pops = prop.make_prop_ops("restricted", "unrestricted")
    else:  # the headline's mixed form
        pops = prop.make_prop_ops("restricted", "unrestricted", precision=precision, taylor_precision=precision,
                                  taylor_hi_from=3)
    params = prop.QmcParams(dt=dt, n_walkers=W, seed=100 + i)
    pctx, mctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params), mops.build_meas_ctx(ham, tr)
    w0 = (torch.from_numpy(syn.walkers).cuda(), torch.from_numpy(syn.walkers_b).cuda())
    st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=w0)
    return batch.Fragment(sys=sys_, ham=ham, trial=tr, trial_ops=tops, meas_ops=mops, prop_ops=pops, params=params,
                          prop_ctx=pctx, meas_ctx=mctx, state=st)


def _fragment_block_energy(f) -> float:
    """Weighted mean of Re E_L over a fragment's walkers (its own energy kernels and weights)."""
    from trot_amd.meas import _h0

    eng = f.meas_ops.engine
    _, g = eng.overlap_green(f.state.walkers, f.trial)
    e = eng.energy_from_green(g, _h0(f.ham), f.meas_ctx).real
    w = f.state.weights
    return float((w * e).sum() / w.sum())


def batch_rates(steps: int = 20) -> dict:
    """The batching form: 32 fragments (N ~ U[40, 200] seed 4, n = round(N/5), G = 4N, UHF trial (n, n-1), 128
    walkers each, dt = 0.005, L / sqrt(N)), FP64: one fragment after another with eager launches, one HIP graph
    of per-fragment launches on 4 branches, one HIP graph of grouped launches with the whole propagation of
    every walker set in one fused launch (the default), and the same with the propagation as one-body / copy /
    Taylor-term launches on autotuned tiles; then the same fragments in the headline's mixed form (split-BF16
    GEMMs, Taylor terms 3-6 one bf16 pass) one after another, through the graph of per-fragment launches on
    4 branches and through the graph of grouped launches.
    frac = FP64 MFMA ideal of the step's GEMMs (VHS on the stored half of V, force bias, Taylor, one-body)
    over the grouped step's time."""
    import torch

    from trot_amd import _lib, batch, synthetic

    W, dt = 128, 0.005
    sizes = synthetic.fragment_sizes(32, seed=4)
    out = {"workload": "32 fragments, N ~ U[40, 200], n = N/5, G = 4N, UHF, 128 walkers each", "steps": steps}
    fields = None
    blocks = {}
    for mode in ("eager", "graph_4_branches", "grouped_graph", "grouped_graph_unfused", "mixed_eager",
                 "mixed_graph_4_branches", "mixed_grouped_graph"):
        prec = "bf16x3" if mode.startswith("mixed") else "f64"
        frags = [_fragment(i, N, n, G, W, dt, 1.0 / np.sqrt(N), prec) for i, (N, n, G) in enumerate(sizes)]
        if fields is None:
            fields = [[torch.from_numpy(synthetic.fields(100 + i, s, W, f.n_fields)).cuda() for i, f in
                       enumerate(frags)] for s in range(steps + 3)]
            # Taylor tile shapes of every (walkers, N, n), tuned once into the process-wide table that every
            # mode's Taylor launches read
            out["taylor_tiles"] = {f"{k[1]}x{k[2]}": v for k, v in
                                   sorted(batch.FragmentBatch(frags, grouped=True).autotune().items())}
        if mode in ("eager", "mixed_eager"):
            for s in range(2):
                for i, f in enumerate(frags):
                    f.state = f.step(fields[s][i])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for s in range(2, steps + 2):
                for i, f in enumerate(frags):
                    f.state = f.step(fields[s][i])
            torch.cuda.synchronize()
        else:
            fb = batch.FragmentBatch(frags, grouped="grouped" in mode, fused=mode == "grouped_graph")
            assert fb.grouped == ("grouped" in mode)
            if mode.startswith("grouped"):
                fb.autotune()  # the grouped plan takes the tile shapes of its own batch
            flat = [fb.pack_fields([x.cpu().numpy() for x in fs]) for fs in fields]
            fb.capture(flat[0])
            fb.step(flat[1])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for s in range(2, steps + 2):
                fb.step(flat[s])
            torch.cuda.synchronize()
            del fb
        ms = 1e3 * (time.perf_counter() - t0) / steps
        out[f"{mode}_ms_per_step"] = round(ms, 3)
        log(f"batch {mode}: {ms:.2f} ms per step of all 32 fragments")
        if mode in ("grouped_graph", "mixed_grouped_graph"):
            # every mode ran the same steps with the same fields from the same states: the block energy of
            # each fragment (each form's own energy kernels, its own weights) after them
            blocks[mode] = [_fragment_block_energy(f) for f in frags]
        del frags
        torch.cuda.empty_cache()
    symcols = lambda N: 64 * ((N + 7) // 8) * ((N + 7) // 8 + 1) // 2
    flop = sum(4 * W * G * symcols(N) + 4 * W * G * N * (2 * n - 1) + 8 * W * N * N * (2 * n - 1) * 6
               + 8 * W * N * N * (2 * n - 1) for N, n, G in sizes)
    out["gemm_flop_per_step"] = float(flop)
    out["fp64_t_star_ms"] = round(1e3 * flop / (PEAK["f64"] * 1e12), 3)
    out["grouped_fp64_frac"] = round(out["fp64_t_star_ms"] / out["grouped_graph_ms_per_step"], 4)
    out["speedup_grouped_vs_eager"] = round(out["eager_ms_per_step"] / out["grouped_graph_ms_per_step"], 3)
    out["mixed_speedup_vs_fp64_grouped"] = round(out["grouped_graph_ms_per_step"]
                                                 / out["mixed_grouped_graph_ms_per_step"], 3)
    d = [abs(m - f) * 1e3 for m, f in zip(blocks["mixed_grouped_graph"], blocks["grouped_graph"])]
    out["mixed_vs_fp64_block_energy_mHa"] = {"steps": steps + 2, "max_over_fragments": round(max(d), 6),
                                             "mean_over_fragments": round(float(np.mean(d)), 6)}
    return out


def differentiation_rates(steps: int = 40, every: int = 20) -> dict:
    """The differentiation form at its sizes (N=128, n=16, G=640, 512 walkers, dt = 0.005, seed 3, L / sqrt(N)):
    ms per step of the product's FP64 step, of the tangent propagation (trot_amd.diff, one direction), and of
    the reverse-mode adjoint (trot_amd.adjoint: forward with checkpoints every ``every`` steps, recomputation,
    reverse; the full dE/dh1), FP64 and FP32 Cholesky contractions."""
    import torch

    from trot_amd import adjoint, diff, meas, prop, synthetic, trial
    from trot_amd.core import HamChol, System

    N, n, G, W, dt = 128, 16, 640, 512, 0.005
    syn = synthetic.make_synthetic(N, n, G, W, 3, 256.0, chol_scale=1.0 / np.sqrt(N))
    rng = np.random.default_rng(33)
    a = rng.standard_normal((N, N))
    D = 0.05 * (np.triu(a) + np.triu(a, 1).T)
    params = prop.QmcParams(dt=dt, n_walkers=W, seed=3, n_prop_steps=steps)
    w0 = torch.from_numpy(syn.walkers).cuda()
    fields = [torch.from_numpy(synthetic.fields(3, k, W, G)).cuda() for k in range(steps)]
    out = {"workload": "n128_o16_g640_w512", "steps": steps, "checkpoint_every": every}

    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(syn.h0, syn.h1, syn.chol).to_device()
    tr = trial.RhfTrial(syn.trial)
    mops, tops = meas.make_rhf_meas_ops(sys_), trial.make_rhf_trial_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted")
    pctx, mctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params), mops.build_meas_ctx(ham, tr)
    st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops, params=params,
                              initial_walkers=w0)
    kw = dict(params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
              prop_ctx=pctx)
    st = pops.step(st, fields=fields[0], **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        st = pops.step(st, fields=fields[k], **kw)
    torch.cuda.synchronize()
    out["primal_fp64_ms_per_step"] = round(1e3 * (time.perf_counter() - t0) / steps, 3)
    del st

    tc = diff.build_ctx(syn.h0, syn.h1, syn.chol, syn.trial, D, 0.0, params)
    diff.run(tc, w0, lambda k: fields[k], 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res_t = diff.run(tc, w0, lambda k: fields[k], steps)
    torch.cuda.synchronize()
    out["tangent_ms_per_step"] = round(1e3 * (time.perf_counter() - t0) / steps, 3)
    for prec in ("f64", "f32"):
        ac = adjoint.build_ctx(syn.h0, syn.h1, syn.chol, syn.trial, params, precision=prec)
        adjoint.run(ac, w0, lambda k: fields[k], 2, checkpoint_every=every)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = adjoint.run(ac, w0, lambda k: fields[k], steps, checkpoint_every=every)
        torch.cuda.synchronize()
        out[f"adjoint_{prec}_ms_per_step"] = round(1e3 * (time.perf_counter() - t0) / steps, 3)
        out[f"adjoint_{prec}_dE_dlambda"] = float(np.sum(res.grad_h1.numpy() * D))
    out["tangent_dE_dlambda"] = res_t.denergy
    log(f"differentiation: {out}")
    return out


def cisd_rates(args) -> dict:
    """CISD trial (trot_amd.cisd) on its own workload: step rate and local-energy time, FP64
    (mixed_precision=False, testing=True) and the reference's default precision (mixed_precision=True,
    testing=False: split-BF16 force bias and doubles term) with the mixed propagation of the main line;
    the two energies compared on the same walkers."""
    import torch

    from trot_amd import cisd, prop, synthetic
    from trot_amd.core import HamChol, System

    cfg = synthetic.WORKLOADS[args.cisd_workload]
    N, n, G, W, dt = cfg["norb"], cfg["nocc"], cfg["nchol"], cfg["n_walkers"], cfg["dt"]
    dev = torch.device("cuda")
    h0, h1, chol, ci1, ci2, w0 = synthetic.make_synthetic_cisd(N, n, G, W, cfg["seed"], cfg["tau"], dev)
    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(h0, h1, chol)
    tr = cisd.CisdTrial(ci1, ci2)
    tops = cisd.make_cisd_trial_ops(sys_)
    out = {"workload": args.cisd_workload, "nbasis": N, "nocc": n, "nvir": N - n, "nchol": G, "n_walkers": W,
           "data": "synthetic (make_synthetic_cisd)"}
    energies = {}
    state = None
    for tag, mixed in (("fp64", False), ("mixed", True)):
        mops = cisd.make_cisd_meas_ops(sys_, mixed_precision=mixed, testing=not mixed)
        if mixed:
            pops = prop.make_prop_ops("restricted", "restricted", precision=args.precision,
                                      taylor_precision=args.taylor_precision, taylor_hi_from=args.taylor_hi_from)
        else:
            pops = prop.make_prop_ops("restricted", "restricted")
        params = prop.QmcParams(dt=dt, n_walkers=W, seed=cfg["seed"])
        pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
        mctx = mops.build_meas_ctx(ham, tr)
        st = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                  params=params, initial_walkers=w0)
        kw = dict(params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops, meas_ctx=mctx,
                  prop_ctx=pctx)
        for _ in range(2):
            st = pops.step(st, **kw)
        torch.cuda.synchronize()
        steps = 5
        t0 = time.perf_counter()
        for _ in range(steps):
            st = pops.step(st, **kw)
        torch.cuda.synchronize()
        t_step = (time.perf_counter() - t0) / steps
        if state is None:
            state = st.walkers.clone()  # both precisions measure the energy on the FP64 run's walkers
        _, greens = mops.engine.overlap_green(state, tr)
        mops.engine.energy_from_green(greens, h0, mctx)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e = mops.engine.energy_from_green(greens, h0, mctx)
        torch.cuda.synchronize()
        t_e = time.perf_counter() - t0
        energies[tag] = e.cpu().numpy()
        out[f"{tag}_ms_per_step"] = round(1e3 * t_step, 3)
        out[f"{tag}_walker_steps_per_s"] = round(W / t_step, 1)
        out[f"{tag}_energy_ms"] = round(1e3 * t_e, 2)
        log(f"cisd {tag}: {1e3 * t_step:.2f} ms/step, energy {1e3 * t_e:.1f} ms")
        del st, pctx, mctx, pops, mops, greens
    d = energies["mixed"] - energies["fp64"]
    out["mixed_vs_fp64_energy"] = {"max_abs_Ha": float(np.max(np.abs(d))),
                                   "mean_diff_Ha": float(abs(np.mean(d.real))),
                                   "mean_fp64_Ha": float(np.mean(energies["fp64"].real))}
    out["speedup_mixed_vs_fp64"] = round(out["fp64_ms_per_step"] / out["mixed_ms_per_step"], 3)
    return out


def cpu_baseline(syn, r: Runner, args, dt):
    """NumPy FP64 oracle (oracle/afqmc_oracle.py) timed on a bounded sample of the same workload."""
    from oracle import afqmc_oracle as O

    try:
        from threadpoolctl import threadpool_info

        threads = max((int(i.get("num_threads", 1)) for i in threadpool_info()), default=1)
    except Exception:  # pragma: no cover
        threads = os.cpu_count() or 1
    W = min(args.cpu_walkers, syn.walkers.shape[0])
    N = r.pctx.norb
    chol_flat = syn.chol.reshape(syn.chol.shape[0], -1).cpu().numpy()
    ctx = O.PropCtx(dt=dt, sqrt_dt=float(np.sqrt(dt)), exp_h1_half=r.pctx.exp_h1_half.cpu().numpy(),
                    mf_shifts=r.pctx.mf_shifts.cpu().numpy(), h0_prop=r.pctx.h0_prop, chol_flat=chol_flat, norb=N)
    c = syn.trial
    rc = r.mctx.rot_chol.cpu().numpy()
    mctx = O.RhfMeasCtx(rot_h1=r.mctx.rot_h1.cpu().numpy(), rot_chol=rc, rot_chol_flat=rc.reshape(rc.shape[0], -1))
    ops = O.RhfOps(overlap=lambda w: O.overlap_r(w, c), force_bias=lambda w: O.force_bias_rw(w, mctx, c),
                   energy=None)
    w0 = syn.walkers[:W].cpu().numpy()
    e_est = float(r.init().e_estimate.item())  # walker 0's energy, as the GPU state starts
    st = O.PropState(walkers=w0, weights=np.ones(W), overlaps=ops.overlap(w0), pop_control_ene_shift=e_est,
                     e_estimate=e_est)
    params = O.QmcParams(dt=dt, n_walkers=W)
    rng = np.random.default_rng(0)
    G = chol_flat.shape[0]
    log(f"cpu baseline: {W} walkers, warm-up step ({threads} BLAS threads)")
    st = O.afqmc_step(st, rng.standard_normal((W, G)), ctx, ops, params)
    t0 = time.perf_counter()
    for s in range(args.cpu_steps):
        st = O.afqmc_step(st, rng.standard_normal((W, G)), ctx, ops, params)
        log(f"cpu baseline: step {s + 1}/{args.cpu_steps} at {time.perf_counter() - t0:.1f}s")
    dt_cpu = time.perf_counter() - t0
    rate = W * args.cpu_steps / dt_cpu
    log(f"cpu baseline: {rate:.2f} walker-steps/s, weights in [{st.weights.min():.4f}, {st.weights.max():.4f}]")
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        affinity = None
    return {"value": round(rate, 3), "unit": "walker-steps/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(), "affinity_cpus": affinity,
            "sample": f"{W} of the workload's {syn.walkers.shape[0]} walkers x {args.cpu_steps} FP64 steps after 1 "
                      f"warm-up step, numpy oracle (oracle/afqmc_oracle.py) with {threads} BLAS threads, same "
                      f"(scaled) Hamiltonian and trial, {dt_cpu:.1f} s"}


if __name__ == "__main__":
    main()
