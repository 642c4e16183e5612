"""Write tests/golden/config2_scaled_100steps.npz: the CPU oracle's 100 FP64 steps of the scale-up case.

Workload n256_o32_g1024_w1024 (N = 256, n = 32, G = 1024, 1024 walkers, dt = 0.01, seed 1, tau 256) with
every L_gamma scaled by 1/sqrt(N), so that the walkers keep O(1) weights for all 100 steps (the recipe as
written overflows within a few steps).  Fields: trot_amd.synthetic.fields(1, step, 1024, 1024).

Stored (all from oracle/afqmc_oracle.py, the restatement of trot/prop/afqmc.py:79-144):
  pop_shift[s], nodes[s]          after each step s = 1..100;
  weights_k, overlaps_k           all walkers after steps k = 1, 10, 50, 100;
  probes_k (W, 4) complex         r_j^T w c_j for four fixed random vectors (r_j, c_j) per walker: a random
                                  linear functional of every walker matrix, so an error in any element shows;
  fro_k (W,)                      the Frobenius norm of each walker.
The probe vectors are drawn from default_rng(2024) and stored too.

Run from the repository root: python tests/golden/make_config2_golden.py   (about 2-4 minutes on 8 cores)
"""

from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import afqmc_oracle as O  # noqa: E402
from trot_amd import synthetic  # noqa: E402

CHECKPOINTS = (1, 10, 50, 100)
OUT = Path(__file__).resolve().parent / "config2_scaled_100steps.npz"


def main() -> None:
    cfg = synthetic.WORKLOADS["n256_o32_g1024_w1024"]
    N, n, G, W, dt = cfg["norb"], cfg["nocc"], cfg["nchol"], cfg["n_walkers"], cfg["dt"]
    syn = synthetic.make_synthetic(N, n, G, W, cfg["seed"], cfg["tau"], chol_scale=1.0 / np.sqrt(N))
    ops = O.make_rhf_ops(syn.h0, syn.h1, syn.chol, syn.trial)
    ctx = O.build_prop_ctx(syn.h0, syn.h1, syn.chol, O.rhf_rdm1(syn.trial), dt)
    params = O.QmcParams(dt=dt, n_walkers=W)
    prng = np.random.default_rng(2024)
    r = prng.standard_normal((4, N)) + 1j * prng.standard_normal((4, N))
    c = prng.standard_normal((4, n)) + 1j * prng.standard_normal((4, n))
    st = O.init_prop_state(syn.walkers, ops, W)
    out = {"probe_r": r, "probe_c": c, "e_estimate": np.float64(st.e_estimate)}
    pops, nodes = [], []
    t0 = time.time()
    for s in range(100):
        st = O.afqmc_step(st, synthetic.fields(cfg["seed"], s, W, G), ctx, ops, params)
        pops.append(st.pop_control_ene_shift)
        nodes.append(st.node_encounters)
        k = s + 1
        if k in CHECKPOINTS:
            out[f"weights_{k}"] = st.weights.copy()
            out[f"overlaps_{k}"] = st.overlaps.copy()
            out[f"probes_{k}"] = np.einsum("jp,wpq,jq->wj", r, st.walkers, c)
            out[f"fro_{k}"] = np.sqrt(np.sum(np.abs(st.walkers) ** 2, axis=(1, 2)))
        print(f"step {k}/100 at {time.time() - t0:.0f}s, pop shift {st.pop_control_ene_shift:.6f}", flush=True)
    out["pop_shift"] = np.array(pops)
    out["nodes"] = np.array(nodes, dtype=np.int64)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e3:.0f} kB)")


if __name__ == "__main__":
    main()
