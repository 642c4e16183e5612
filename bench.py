"""Throughput of the phaseless AFQMC walker step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--precision f64|f32|bf16x3]

Workload (default): the nbasis=512 synthetic system, N=512 orbitals, n=64
occupied, G=2560 Cholesky vectors, 2048 complex128 walkers, dt=0.005, RHF
trial, restricted walkers (random-init synthetic Hamiltonian, see
trot_amd/synthetic.py).  A step is one phaseless step for all walkers: field
draw, force bias, shifted fields, VHS build, one-body / Taylor / one-body
propagation, overlaps and Green's functions, weight update.  All inputs are
resident in HBM before the timed region.

Prints one JSON line: walker-steps/s (whole job), the roofline of the
dominant kernel (the VHS GEMM; duration from HIP events recorded around its
launches inside the timed steps), the block-inclusive rate (energy, QR and
SR every n_prop_steps=50 steps), and a CPU baseline: the NumPy FP64 oracle
timed on a bounded sample of the same workload on the host cores.
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
METRIC = "walker-steps/s at nbasis=512 (phaseless AFQMC step)"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="n512_o64_g2560_w2048")
    ap.add_argument("--precision", default="f64", choices=["f64", "f32", "bf16x3"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-block", action="store_true", help="skip the block-tail timing")
    ap.add_argument("--cpu-walkers", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=2)
    return ap.parse_args()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(kernel_tag: str, workload: str):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary, if any."""
    p = ROOT / "profiles" / f"{kernel_tag}_{workload}_traffic.md"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    args = parse()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU only: this build has no multi-GPU path (--gpus must be 1)")
    import torch

    from trot_amd import _lib, blocks, meas, prop, synthetic, trial
    from trot_amd.core import HamChol, System

    _lib.lib()
    cfg = synthetic.WORKLOADS[args.workload]
    N, n, G, W, dt = cfg["norb"], cfg["nocc"], cfg["nchol"], cfg["n_walkers"], cfg["dt"]
    dev = torch.device("cuda")
    t0 = time.perf_counter()
    syn = synthetic.make_synthetic(N, n, G, W, cfg["seed"], cfg["tau"], device=dev)
    log(f"synthetic system N={N} n={n} G={G} W={W} in {time.perf_counter() - t0:.1f}s")

    sys_ = System(N, (n, n), "restricted")
    ham = HamChol(syn.h0, torch.from_numpy(syn.h1).to(dev), syn.chol)
    tr = trial.RhfTrial(syn.trial)
    tops = trial.make_rhf_trial_ops(sys_)
    mops = meas.make_rhf_meas_ops(sys_)
    pops = prop.make_prop_ops("restricted", "restricted", precision=args.precision)
    params = prop.QmcParams(dt=dt, n_walkers=W, seed=cfg["seed"])
    pctx = pops.build_prop_ctx(ham, tops.get_rdm1(tr), params)
    mctx = mops.build_meas_ctx(ham, tr)
    state = pops.init_prop_state(sys=sys_, ham_data=ham, trial_ops=tops, trial_data=tr, meas_ops=mops,
                                 params=params, initial_walkers=syn.walkers)
    torch.cuda.synchronize()
    log(f"contexts and initial state ready at {time.perf_counter() - t0:.1f}s")

    def step(st):
        return pops.step(st, params=params, ham_data=ham, trial_data=tr, trial_ops=tops, meas_ops=mops,
                         meas_ctx=mctx, prop_ctx=pctx)

    for i in range(args.warmup):
        state = step(state)
        torch.cuda.synchronize()
        log(f"warmup step {i + 1}/{args.warmup} done")

    timer = prop.KernelTimer()
    pctx.low["timer"] = timer
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        state = step(state)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    pctx.low.pop("timer")
    ms_step = 1e3 * t_total / args.steps
    value = W * args.steps / t_total
    log(f"timed {args.steps} steps: {ms_step:.2f} ms/step, {value:.1f} walker-steps/s")

    vhs_ms = timer.mean_ms("vhs")
    vhs_flop = 4.0 * W * G * N * N
    achieved = vhs_flop / (vhs_ms * 1e-3) / 1e12
    peak = PEAK[args.precision]
    traffic = traffic_from_profiles("vhs", args.workload)
    roofline = {"bound": "mfma", "kernel": "VHS GEMM (vhs_build)", "achieved": round(achieved, 2),
                "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
                "launch_ms": round(vhs_ms, 3), "flop_per_launch": vhs_flop}
    step_flop = 4.0 * G * N * N + 4.0 * G * n * N + 8.0 * params.n_exp_terms * N * N * n + 8.0 * N * N * n \
        + 16.0 * n * n * N
    step_tflops = step_flop * W / (ms_step * 1e-3) / 1e12

    block_rate = None
    if not args.no_block:
        torch.cuda.synchronize()
        tb = time.perf_counter()
        reps = 2
        for _ in range(reps):
            wq = __import__("trot_amd.walkers", fromlist=["x"]).orthonormalize(state.walkers, "restricted")
            ov, greens = mops.engine.overlap_green(wq, tr)
            e = mops.engine.energy_from_green(greens, 0.0, mctx)
            from trot_amd import walkers as wk

            idx, wn = wk.sr_indices(state.weights, 0.5)
            _ = wk.gather(wq, idx)
        torch.cuda.synchronize()
        t_tail = (time.perf_counter() - tb) / reps
        nps = params.n_prop_steps
        block_rate = W * nps / (nps * t_total / args.steps + t_tail)
        log(f"block tail (QR + energy + SR): {1e3 * t_tail:.1f} ms; block-inclusive {block_rate:.1f} walker-steps/s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(syn, pctx, mctx, args, dt)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64 (complex128 walkers, FP64 MFMA)", "f32": "f32 Cholesky GEMMs / f64 det, QR, weights",
                  "bf16x3": "split-bf16 Cholesky GEMMs / f64 det, QR, weights"}[args.precision],
        "data": "synthetic (seeded numpy PCG64 Hamiltonian, trial and walkers; on-device N(0,1) fields)",
        "config": {"workload": args.workload, "nbasis": N, "nocc": n, "nchol": G, "n_walkers": W, "dt": dt,
                   "n_exp_terms": params.n_exp_terms, "trial": "RHF", "walker_kind": "restricted",
                   "precision": args.precision},
        "step_tflops": round(step_tflops, 2),
        "block_inclusive_walker_steps_per_s": None if block_rate is None else round(block_rate, 2),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(syn, pctx, mctx, args, dt):
    """NumPy FP64 oracle (oracle/afqmc_oracle.py) timed on a bounded sample of the same workload."""
    import torch

    from oracle import afqmc_oracle as O

    try:
        from threadpoolctl import threadpool_info

        threads = max((int(i.get("num_threads", 1)) for i in threadpool_info()), default=1)
    except Exception:  # pragma: no cover
        threads = os.cpu_count() or 1
    W = min(args.cpu_walkers, syn.walkers.shape[0])
    N = pctx.norb
    chol_flat = syn.chol.reshape(syn.chol.shape[0], -1).cpu().numpy()
    ctx = O.PropCtx(dt=dt, sqrt_dt=float(np.sqrt(dt)), exp_h1_half=pctx.exp_h1_half.cpu().numpy(),
                    mf_shifts=pctx.mf_shifts.cpu().numpy(), h0_prop=pctx.h0_prop, chol_flat=chol_flat, norb=N)
    c = syn.trial
    rot = O.RhfMeasCtx(rot_h1=mctx.rot_h1.cpu().numpy(), rot_chol=mctx.rot_chol.cpu().numpy(),
                       rot_chol_flat=mctx.rot_chol.cpu().numpy().reshape(mctx.rot_chol.shape[0], -1))
    ops = O.RhfOps(overlap=lambda w: O.overlap_r(w, c), force_bias=lambda w: O.force_bias_rw(w, rot, c),
                   energy=None)
    w0 = syn.walkers[:W].cpu().numpy()
    st = O.PropState(walkers=w0, weights=np.ones(W), overlaps=ops.overlap(w0), pop_control_ene_shift=0.0,
                     e_estimate=0.0)
    params = O.QmcParams(dt=dt, n_walkers=W)
    rng = np.random.default_rng(0)
    G = chol_flat.shape[0]
    st = O.afqmc_step(st, rng.standard_normal((W, G)), ctx, ops, params)  # warm-up
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        st = O.afqmc_step(st, rng.standard_normal((W, G)), ctx, ops, params)
    dt_cpu = time.perf_counter() - t0
    rate = W * args.cpu_steps / dt_cpu
    log(f"cpu baseline: {W} walkers x {args.cpu_steps} steps in {dt_cpu:.1f}s -> {rate:.2f} walker-steps/s "
        f"({threads} threads)")
    del torch
    return {"value": round(rate, 3), "unit": "walker-steps/s", "cores": threads, "kind": "port",
            "sample": f"{W} of the workload's walkers x {args.cpu_steps} steps (after 1 warm-up step), "
                      f"numpy FP64 oracle, same Hamiltonian and trial, {dt_cpu:.1f} s"}


if __name__ == "__main__":
    main()
