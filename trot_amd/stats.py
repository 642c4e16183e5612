"""Host block statistics (trot/stat_utils.py:45-236), for block-average checks.

The ratio estimator mu = sum(w e) / sum(w) with a leave-one-out jackknife
standard error per block size B; the reported error is the first plateau of
the SE curve (a simpler plateau rule than the reference's _pick_plateau).
"""

from __future__ import annotations

import numpy as np


def blocking_analysis_ratio(ene, wt, block_grid=None, *, min_blocks: int = 20, print_q: bool = False) -> dict:
    """trot/stat_utils.py:45-217 (ratio estimator, LOO jackknife per block size)."""
    ene = np.asarray(ene, float).ravel()
    wt = np.asarray(wt, float).ravel()
    n = ene.size
    s = wt * ene
    mu = float(s.sum() / wt.sum()) if n else float("nan")
    if block_grid is None:
        raw = np.unique(np.rint(np.geomspace(1, max(2, n // min_blocks), 18)).astype(int))
        block_grid = [int(b) for b in raw if b >= 1 and (n // b) >= min_blocks]
    bs, ses = [], []
    for b in block_grid:
        g = n // b
        if g < 5:
            continue
        u = g * b
        sg = s[:u].reshape(g, b).sum(1)
        ng = wt[:u].reshape(g, b).sum(1)
        st, nt = sg.sum(), ng.sum()
        den = nt - ng
        mu_loo = np.where(np.abs(den) > 1e-18, (st - sg) / np.where(den == 0, 1, den), st / nt)
        var = (g - 1) / g * np.sum((mu_loo - mu_loo.mean()) ** 2)
        bs.append(b)
        ses.append(float(np.sqrt(max(var, 0.0))))
    if not bs:
        return {"mu": mu, "se_star": None, "B_star": None, "Bs": [], "SEs": []}
    se_arr = np.array(ses)
    k = len(se_arr) - 1
    for i in range(1, len(se_arr)):
        if se_arr[i] <= se_arr[i - 1] * 1.03:
            k = i - 1
            break
    if print_q:
        print(f"blocking: mu = {mu:.10f}, SE* = {se_arr[k]:.3e} at B = {bs[k]}")
    return {"mu": mu, "se_star": float(se_arr[k]), "B_star": int(bs[k]), "Bs": bs, "SEs": ses}


def reject_outliers(data, obs: int, m: float = 10.0, min_threshold: float = 1e-5):
    """trot/stat_utils.py:220-236."""
    data = np.asarray(data)
    target = data[:, obs]
    med = np.median(target)
    d = np.abs(target - med)
    mdev = np.median(d)
    q1, q3 = np.percentile(target, [25, 75])
    dispersion = max(mdev, (q3 - q1) / 1.349, min_threshold)
    mask = d / dispersion < m
    return data[mask], mask
