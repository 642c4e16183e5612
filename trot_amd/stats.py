"""Host block statistics (trot/stat_utils.py:12-313).

Block energies come back from the GPU as one scalar pair (energy, weight) per block; the error analysis
over a few thousand blocks is host NumPy work, restated here with the reference's rules so that a run
reports the same mean, the same error bar and the same chosen block size:

* ``blocking_analysis_ratio``: the ratio estimator mu = sum(w e) / sum(w), with the leave-one-out
  jackknife standard error of mu for each block size B of a geometric grid, and the plateau of that
  curve picked by ``_pick_plateau`` (the first B whose SE has risen by ``min_rise`` over SE(B=1) and
  stays flat to ``flat_tol`` over the next ``k`` sizes, else the first B within 5% of the largest SE);
* ``reject_outliers``: median / MAD / IQR filter on one column of the (energy, weight) table;
* ``rebin_observable`` and ``jackknife_ratios``: the error bar of an array-valued block observable at
  the energy's B*.
"""

from __future__ import annotations

from typing import Any, Dict, Iterable

import numpy as np


def _pick_plateau(block_sizes, se_curve, n_blocks, *, min_blocks: int = 20, min_rise: float = 0.20,
                  flat_tol: float = 0.03, k: int = 3) -> tuple[int, float, int]:
    """The plateau of the SE-vs-block-size curve (trot/stat_utils.py:12-42).

    Only sizes that leave at least ``min_blocks`` blocks count.  Among them, the first size whose SE is at
    least (1 + min_rise) SE(first) and whose next ``k`` steps each change the SE by at most
    ``flat_tol`` of the step's start wins.  With none, the first size within 5% of the largest finite
    SE; with no finite SE at all, or no size with enough blocks, the first size of the curve.
    Returns (B*, SE*, number of blocks at B*).
    """
    bs = np.asarray(block_sizes)
    se = np.asarray(se_curve, dtype=float)
    gs = np.asarray(n_blocks)
    if bs.size == 0:
        raise ValueError("empty blocking curve")
    keep = gs >= min_blocks
    if not keep.any():
        return int(bs[0]), float(se[0]), int(gs[0])
    b2, s2, g2 = bs[keep], se[keep], gs[keep]
    risen = s2 >= (1.0 + min_rise) * s2[0]
    for i in range(b2.size - k):
        if not risen[i]:
            continue
        win = s2[i:i + k + 1]
        if (np.abs(win[1:] - win[:-1]) <= flat_tol * win[:-1]).all():
            return int(b2[i]), float(s2[i]), int(g2[i])
    fin = np.isfinite(s2)
    if not fin.any():
        return int(bs[0]), float(se[0]), int(gs[0])
    top = float(np.max(s2[fin]))
    j = int(np.flatnonzero(s2 >= 0.95 * top)[0])
    return int(b2[j]), float(s2[j]), int(g2[j])


def default_block_grid(n: int, min_blocks: int = 20) -> list[int]:
    """The block sizes tried for n samples (trot/stat_utils.py:71-75): 18 geometric steps from 1 to
    max(2, n // min_blocks), rounded and de-duplicated, kept where they leave min_blocks blocks; the
    largest rounded step is added when it still leaves 5 blocks."""
    raw = np.unique(np.rint(np.geomspace(1, max(2, n // min_blocks), 18)).astype(int))
    grid = [int(b) for b in raw if b >= 1 and n // b >= min_blocks]
    top = int(raw[-1])
    if n // top >= 5 and top not in grid:
        grid.append(top)
    return grid


def _loo_curve(s: np.ndarray, w: np.ndarray, b: int):
    """Leave-one-block-out ratios for block size b: (mu_loo per block, their mean, SE, number of blocks)."""
    g = s.size // b
    u = g * b
    sg = s[:u].reshape(g, b).sum(axis=1)
    wg = w[:u].reshape(g, b).sum(axis=1)
    st, wt = sg.sum(), wg.sum()
    rest = wt - wg
    ok = np.abs(rest) > 1e-18
    with np.errstate(divide="ignore", invalid="ignore"):
        mu_loo = np.where(ok, (st - sg) / np.where(ok, rest, 1.0), st / wt)
    mean = mu_loo.mean()
    var = (g - 1) / g * float(np.sum((mu_loo - mean) ** 2))
    return mu_loo, mean, float(np.sqrt(max(var, 0.0))), g


def blocking_analysis_ratio(ene, wt, block_grid: Iterable[int] | None = None, *, min_blocks: int = 20,
                            min_rise: float = 0.20, flat_tol: float = 0.03, k: int = 3,
                            print_q: bool = False, exact: float | None = None) -> Dict[str, Any]:
    """Blocking analysis of mu = sum(wt ene) / sum(wt) (trot/stat_utils.py:45-217; no plotting).

    Keys: ``mu``, ``block_sizes``, ``se_curve``, ``n_blocks``, ``B_star``, ``se_star``, ``ci95_star``,
    ``estimator_scale_samples`` (the LOO ratios at B* rescaled to the estimator's spread), ``bias`` and
    ``z_score`` (against ``exact`` when given and SE* > 0).  With fewer than 5 blocks at every size the
    curve keys are None.
    """
    e = np.asarray(ene, dtype=float).ravel()
    w = np.asarray(wt, dtype=float).ravel()
    if e.size != w.size:
        raise ValueError(f"{e.size} energies against {w.size} weights")
    n = e.size
    s = w * e
    mu = float(s.sum() / w.sum())
    grid = default_block_grid(n, min_blocks) if block_grid is None else [int(b) for b in block_grid]

    bs, ses, gs, loo = [], [], [], {}
    for b in grid:
        if n // b < 5:
            continue
        mu_loo, mean, se, g = _loo_curve(s, w, b)
        bs.append(b)
        ses.append(se)
        gs.append(g)
        loo[b] = (mu_loo, mean, g)
    if not bs:
        return {"mu": mu, "block_sizes": None, "se_curve": None, "n_blocks": None, "B_star": None,
                "se_star": None, "ci95_star": (None, None), "estimator_scale_samples": None, "bias": None,
                "z_score": None}

    b_arr, se_arr, g_arr = np.array(bs, int), np.array(ses, float), np.array(gs, int)
    b_star, se_star, _ = _pick_plateau(b_arr, se_arr, g_arr, min_blocks=min_blocks, min_rise=min_rise,
                                       flat_tol=flat_tol, k=k)
    mu_loo, mean, g = loo[b_star]
    est = mu + (g - 1) / np.sqrt(g) * (mu_loo - mean)
    bias = z = None
    if exact is not None and np.isfinite(se_star) and se_star > 0:
        bias = mu - float(exact)
        z = bias / se_star
    out = {"mu": mu, "block_sizes": b_arr, "se_curve": se_arr, "n_blocks": g_arr, "B_star": int(b_star),
           "se_star": float(se_star), "ci95_star": (mu - 1.96 * se_star, mu + 1.96 * se_star),
           "estimator_scale_samples": est, "bias": bias, "z_score": z}
    if print_q:
        print(f"mu: {mu:.16g}  SE*: {se_star:.16g}  95% CI: {out['ci95_star']}")
        if z is not None:
            print(f"bias: {bias:.16g}  z: {z:.6g}")
        print("\nBlocking SE curve (ratio LOO):")
        print(f" {'B':>6s} {'G':>6s} {'SE':>14s} {'SE/SE(B=1)':>12s}")
        se0 = float(se_arr[0])
        for b, gg, se in zip(bs, gs, ses):
            r = se / se0 if se0 > 0 and np.isfinite(se0) else float("nan")
            print(f"{'*' if b == b_star else ' '}{b:6d} {gg:6d} {se:14.6e} {r:12.3f}")
        print("")
    return out


def reject_outliers(data, obs: int, m: float = 10.0, min_threshold: float = 1e-5):
    """Rows of ``data`` whose column ``obs`` lies within m dispersions of its median
    (trot/stat_utils.py:220-236); the dispersion is the largest of the median absolute deviation, the
    normal-scaled IQR and ``min_threshold``.  Returns (kept rows, mask)."""
    data = np.asarray(data)
    col = data[:, obs]
    dev = np.abs(col - np.median(col))
    q1, q3 = np.percentile(col, [25, 75])
    disp = max(float(np.median(dev)), float(q3 - q1) / 1.349, min_threshold)
    mask = dev / disp < m
    return data[mask], mask


def jackknife_ratios(num, denom) -> tuple[np.ndarray, np.ndarray]:
    """Leave-one-out jackknife of sum(num) / sum(denom) for array-valued samples
    (trot/stat_utils.py:239-275): num (n, *shape), denom (n,); returns the mean of the n LOO ratios (real
    part) and sqrt((n - 1) var) of them, each of shape ``shape``."""
    num = np.asarray(num)
    den = np.asarray(denom).ravel()
    n = num.shape[0]
    if den.shape[0] != n:
        raise ValueError(f"{n} numerator samples against {den.shape[0]} denominators")
    loo_num = (num.sum(axis=0) - num) / (n - 1)
    loo_den = ((den.sum() - den) / (n - 1)).reshape((n,) + (1,) * (num.ndim - 1))
    ratio = (loo_num / loo_den).real
    return ratio.mean(axis=0), np.sqrt((n - 1) * ratio.var(axis=0))


def rebin_observable(obs, weights, block_size: int) -> tuple[np.ndarray, np.ndarray]:
    """Super-blocks of ``block_size`` consecutive blocks (trot/stat_utils.py:278-313): numerators
    sum(w_b obs_b) of shape (n // block_size, *shape) and denominators sum(w_b); a ragged tail is
    dropped."""
    obs = np.asarray(obs)
    w = np.asarray(weights).ravel()
    g = obs.shape[0] // block_size
    u = g * block_size
    wg = w[:u].reshape(g, block_size)
    og = obs[:u].reshape((g, block_size) + obs.shape[1:])
    num = (wg.reshape((g, block_size) + (1,) * (obs.ndim - 1)) * og).sum(axis=1)
    return num, wg.sum(axis=1)
