"""CPU FP64 oracle for the phaseless AFQMC walker step and local energy.

TEST INFRASTRUCTURE ONLY.  This module is the test oracle, never the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``trot_amd``) never imports,
calls or falls back to anything here.

It restates, in plain NumPy complex128 / float64, the algorithm of the
reference package ``trot`` (JAX).  Every function cites the reference
file:line it follows (paths relative to the reference root).  The reference
itself cannot be imported here (JAX, jaxlib and PySCF are absent), so this
restatement is pinned by closed forms, known-answer tests taken from the
reference's own tests, brute-force Fock-space evaluation on tiny systems, and,
for the field shifts and importance weights, the unbiasedness of importance
sampling (Gauss-Hermite quadrature) and the Fock-space propagator
exp(-dt (H - E_shift)) (see ``oracle/fock.py`` and ``tests/test_oracle_*.py``).

Walker layouts follow the reference: restricted walkers are complex arrays of
shape (W, N, n); unrestricted walkers are pairs ``(wu (W,N,na), wd (W,N,nb))``.
All functions are batched over the leading walker axis (the reference maps
per-walker kernels with ``vmap``; the arithmetic per walker is identical).
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Any, Callable

import numpy as np
import scipy.linalg

# ---------------------------------------------------------------------------
# parameters and state (trot/prop/types.py:20-47)
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class QmcParams:
    """trot/prop/types.py:30-47 (defaults identical)."""

    dt: float = 0.005
    n_chunks: int = 1
    n_exp_terms: int = 6
    n_prop_steps: int = 50
    n_blocks: int = 200
    n_walkers: int = 200
    seed: int = 0
    pop_control_damping: float = 0.1
    weight_floor: float = 1.0e-3
    weight_cap: float = 100.0
    shift_ema: float = 0.1
    n_eql_blocks: int = 20


@dataclass
class PropState:
    """trot/prop/types.py:20-27 (rng_key replaced by explicit field/zeta inputs)."""

    walkers: Any
    weights: np.ndarray
    overlaps: np.ndarray
    pop_control_ene_shift: float
    e_estimate: float
    node_encounters: int = 0

    def copy(self) -> "PropState":
        w = self.walkers
        w = (w[0].copy(), w[1].copy()) if isinstance(w, tuple) else w.copy()
        return PropState(
            walkers=w,
            weights=self.weights.copy(),
            overlaps=self.overlaps.copy(),
            pop_control_ene_shift=float(np.real(self.pop_control_ene_shift)),
            e_estimate=float(np.real(self.e_estimate)),
            node_encounters=int(self.node_encounters),
        )


# ---------------------------------------------------------------------------
# propagation context (trot/prop/chol_afqmc_ops.py:74-127)
# ---------------------------------------------------------------------------


@dataclass
class PropCtx:
    """trot/prop/chol_afqmc_ops.py:16-50 (CholAfqmcCtx)."""

    dt: float
    sqrt_dt: float
    exp_h1_half: np.ndarray
    mf_shifts: np.ndarray
    h0_prop: complex
    chol_flat: np.ndarray
    norb: int


def total_rdm1(rdm1: np.ndarray) -> np.ndarray:
    """trot/prop/chol_afqmc_ops.py:57-60: a (2,N,N) spin-block rdm1 is summed."""
    rdm1 = np.asarray(rdm1)
    if rdm1.ndim == 3 and rdm1.shape[0] == 2:
        return rdm1[0] + rdm1[1]
    return rdm1


def mf_shifts(chol: np.ndarray, rdm1: np.ndarray) -> np.ndarray:
    """trot/prop/chol_afqmc_ops.py:74-76: mf_g = i * sum_ij L_g,ij dm_ji."""
    dm = total_rdm1(rdm1)
    return 1.0j * np.einsum("gij,ji->g", chol, dm, optimize=True)


def h1_eff(h1: np.ndarray, chol: np.ndarray, mf: np.ndarray) -> np.ndarray:
    """trot/prop/chol_afqmc_ops.py:90-100: h1 - 1/2 sum_g L_g L_g - sum_g Re(i mf_g) L_g."""
    v0m = 0.5 * np.einsum("gik,gkj->ij", chol, chol, optimize=True)
    mf_r = (1.0j * mf).real
    v1m = np.einsum("g,gik->ik", mf_r, chol, optimize=True)
    return h1 - v0m - v1m


def build_prop_ctx(h0: float, h1: np.ndarray, chol: np.ndarray, rdm1: np.ndarray, dt: float) -> PropCtx:
    """trot/prop/chol_afqmc_ops.py:103-127 (_build_prop_ctx), FP64 chol_flat."""
    mf = mf_shifts(chol, rdm1)
    h0_prop = -h0 - 0.5 * np.sum(mf**2)
    he = h1_eff(h1, chol, mf)
    exp_h1_half = scipy.linalg.expm(-0.5 * dt * he)  # chol_afqmc_ops.py:79-80
    chol_flat = chol.reshape(chol.shape[0], -1).astype(np.float64)
    return PropCtx(
        dt=float(dt),
        sqrt_dt=float(np.sqrt(dt)),
        exp_h1_half=exp_h1_half,
        mf_shifts=mf,
        h0_prop=complex(h0_prop),
        chol_flat=chol_flat,
        norb=chol.shape[1],
    )


# ---------------------------------------------------------------------------
# trials (trot/trial/rhf.py, trot/trial/uhf.py)
# ---------------------------------------------------------------------------


def rhf_rdm1(c: np.ndarray) -> np.ndarray:
    """trot/trial/rhf.py:39-42."""
    dm = c @ c.conj().T
    return np.stack([dm, dm], axis=0)


def uhf_rdm1(ca: np.ndarray, cb: np.ndarray) -> np.ndarray:
    """trot/trial/uhf.py:40-45."""
    return np.stack([ca @ ca.conj().T, cb @ cb.conj().T], axis=0)


def overlap_r(w: np.ndarray, c: np.ndarray) -> np.ndarray:
    """trot/trial/rhf.py:45-47: det(C^H w)^2, batched over walkers."""
    m = np.einsum("ip,wiq->wpq", c.conj(), w, optimize=True)
    return np.linalg.det(m) ** 2


def overlap_u(w: tuple, ca: np.ndarray, cb: np.ndarray | None = None) -> np.ndarray:
    """trot/trial/rhf.py:50-54 (cb=None: RHF trial) and trot/trial/uhf.py:54-58."""
    wu, wd = w
    cb = ca if cb is None else cb
    mu = np.einsum("ip,wiq->wpq", ca.conj(), wu, optimize=True)
    md = np.einsum("ip,wiq->wpq", cb.conj(), wd, optimize=True)
    return np.linalg.det(mu) * np.linalg.det(md)


# ---------------------------------------------------------------------------
# measurement context and kernels (trot/meas/rhf.py, trot/meas/uhf.py)
# ---------------------------------------------------------------------------


@dataclass
class RhfMeasCtx:
    """trot/meas/rhf.py:282-305."""

    rot_h1: np.ndarray  # (n, N)
    rot_chol: np.ndarray  # (G, n, N)
    rot_chol_flat: np.ndarray  # (G, n*N)
    memory_mode: str = "high"


@dataclass
class UhfMeasCtx:
    """trot/meas/uhf.py:277-315."""

    rot_h1_a: np.ndarray
    rot_h1_b: np.ndarray
    rot_chol_a: np.ndarray
    rot_chol_b: np.ndarray


def build_meas_ctx_rhf(h1: np.ndarray, chol: np.ndarray, c: np.ndarray, memory_mode: str = "high") -> RhfMeasCtx:
    """trot/meas/rhf.py:308-319."""
    ch = c.conj().T
    rot_h1 = ch @ h1
    rot_chol = np.einsum("pi,gij->gpj", ch, chol, optimize=True)
    return RhfMeasCtx(rot_h1, rot_chol, rot_chol.reshape(rot_chol.shape[0], -1), memory_mode)


def build_meas_ctx_uhf(h1: np.ndarray, chol: np.ndarray, ca: np.ndarray, cb: np.ndarray) -> UhfMeasCtx:
    """trot/meas/uhf.py:318-336."""
    caH, cbH = ca.conj().T, cb.conj().T
    return UhfMeasCtx(
        rot_h1_a=caH @ h1,
        rot_h1_b=cbH @ h1,
        rot_chol_a=np.einsum("pi,gij->gpj", caH, chol, optimize=True),
        rot_chol_b=np.einsum("pi,gij->gpj", cbH, chol, optimize=True),
    )


def half_green(w: np.ndarray, c: np.ndarray) -> np.ndarray:
    """trot/meas/rhf.py:21-25 with m = C^H w (rhf.py:121): G_half = solve(m^T, w^T), (W, n, N)."""
    m = np.einsum("ip,wiq->wpq", c.conj(), w, optimize=True)
    return np.linalg.solve(np.swapaxes(m, 1, 2), np.swapaxes(w, 1, 2))


def rdm1_rw(w: np.ndarray, c: np.ndarray) -> np.ndarray:
    """trot/meas/rhf.py:191-200: g = G_half^T C^H, the same for both spins; (W, 2, N, N)."""
    g = np.swapaxes(half_green(w, c), 1, 2) @ c.conj().T
    return np.stack([g, g], axis=1)


def rdm1_uw(w: tuple, ca: np.ndarray, cb: np.ndarray | None = None) -> np.ndarray:
    """trot/meas/rhf.py:203-216 (cb = ca), trot/meas/uhf.py:97-110; (W, 2, N, N)."""
    cb = ca if cb is None else cb
    ga = np.swapaxes(half_green(w[0], ca), 1, 2) @ ca.conj().T
    gb = np.swapaxes(half_green(w[1], cb), 1, 2) @ cb.conj().T
    return np.stack([ga, gb], axis=1)


def density_corr_from_rdm1(dm: np.ndarray) -> np.ndarray:
    """trot/meas/uhf.py:129-144 per walker: (W, 3, N, N) = (uu, ud, dd) from dm (W, 2, N, N)."""
    ga, gb = dm[:, 0], dm[:, 1]
    na = np.diagonal(ga, axis1=1, axis2=2)
    nb = np.diagonal(gb, axis1=1, axis2=2)
    eye = np.eye(ga.shape[1])
    uu = na[:, :, None] * na[:, None, :] - ga * np.swapaxes(ga, 1, 2) + na[:, :, None] * eye
    dd = nb[:, :, None] * nb[:, None, :] - gb * np.swapaxes(gb, 1, 2) + nb[:, :, None] * eye
    ud = na[:, :, None] * nb[:, None, :]
    return np.stack([uu, ud, dd], axis=1)


def weighted_block_mean(samples: np.ndarray, weights: np.ndarray) -> np.ndarray:
    """trot/prop/blocks.py:259-268: sum_w weights_w samples_w / sum_w weights_w, 0 when the sum is 0."""
    w_sum = np.sum(weights)
    num = np.tensordot(weights, samples, axes=(0, 0))
    return np.zeros_like(num) if w_sum == 0 else num / w_sum


def force_bias_rw(w: np.ndarray, ctx: RhfMeasCtx, c: np.ndarray) -> np.ndarray:
    """trot/meas/rhf.py:119-125: 2 * rot_chol_flat @ vec(G_half)."""
    g = half_green(w, c)
    return 2.0 * np.einsum("gk,wk->wg", ctx.rot_chol_flat, g.reshape(g.shape[0], -1), optimize=True)


def force_bias_uw(w: tuple, ctx: RhfMeasCtx, c: np.ndarray) -> np.ndarray:
    """trot/meas/rhf.py:141-153 (RHF trial, unrestricted walkers)."""
    gu = half_green(w[0], c)
    gd = half_green(w[1], c)
    g = gu + gd
    return np.einsum("gk,wk->wg", ctx.rot_chol_flat, g.reshape(g.shape[0], -1), optimize=True)


def force_bias_uhf(w: tuple, ctx: UhfMeasCtx, ca: np.ndarray, cb: np.ndarray) -> np.ndarray:
    """trot/meas/uhf.py:43-57 (UHF trial, unrestricted walkers)."""
    gu = half_green(w[0], ca)
    gd = half_green(w[1], cb)
    return (np.einsum("gij,wij->wg", ctx.rot_chol_a, gu, optimize=True)
            + np.einsum("gij,wij->wg", ctx.rot_chol_b, gd, optimize=True))


def _exchange_restricted_high(g: np.ndarray, rot_chol: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """trot/meas/rhf.py:57-60: f_g = rot_chol[g] @ G^T; returns (sum_g c_g^2, sum_g sum f*f^T)."""
    f = np.einsum("gij,wkj->wgik", rot_chol, g, optimize=True)  # (W, G, n, n)
    c = np.einsum("wgii->wg", f, optimize=True)
    exc = np.einsum("wgij,wgji->w", f, f, optimize=True)
    return np.sum(c * c, axis=1), exc


def energy_rw(w: np.ndarray, h0: float, ctx: RhfMeasCtx, c: np.ndarray) -> np.ndarray:
    """trot/meas/rhf.py:128-138 with _two_body_energy_restricted (rhf.py:41-60)."""
    g = half_green(w, c)
    e1 = 2.0 * np.einsum("wij,ij->w", g, ctx.rot_h1, optimize=True)
    if ctx.memory_mode == "low":
        c2 = np.zeros(g.shape[0], dtype=complex)
        exc = np.zeros(g.shape[0], dtype=complex)
        for rc in ctx.rot_chol:  # rhf.py:43-55 scan over Cholesky vectors
            f = np.einsum("ij,wkj->wik", rc, g, optimize=True)
            cg = np.einsum("wii->w", f, optimize=True)
            c2 = c2 + cg * cg
            exc = exc + np.einsum("wij,wji->w", f, f, optimize=True)
    else:
        c2, exc = _exchange_restricted_high(g, ctx.rot_chol)
    return h0 + e1 + 2.0 * c2 - exc


def energy_uw(w: tuple, h0: float, ctx: RhfMeasCtx, c: np.ndarray) -> np.ndarray:
    """trot/meas/rhf.py:156-172 with _two_body_energy_unrestricted (rhf.py:63-98)."""
    gu = half_green(w[0], c)
    gd = half_green(w[1], c)
    e1 = np.einsum("wij,ij->w", gu + gd, ctx.rot_h1, optimize=True)
    fu = np.einsum("gij,wkj->wgik", ctx.rot_chol, gu, optimize=True)
    fd = np.einsum("gij,wkj->wgik", ctx.rot_chol, gd, optimize=True)
    cu = np.einsum("wgii->wg", fu, optimize=True)
    cd = np.einsum("wgii->wg", fd, optimize=True)
    exu = np.einsum("wgij,wgji->w", fu, fu, optimize=True)
    exd = np.einsum("wgij,wgji->w", fd, fd, optimize=True)
    e2 = (np.sum(cu * cu, 1) + np.sum(cd * cd, 1) + 2.0 * np.sum(cu * cd, 1) - exu - exd) / 2.0
    return h0 + e1 + e2


def energy_uhf(w: tuple, h0: float, ctx: UhfMeasCtx, ca: np.ndarray, cb: np.ndarray) -> np.ndarray:
    """trot/meas/uhf.py:205-231 (UHF trial, unrestricted walkers)."""
    gu = half_green(w[0], ca)
    gd = half_green(w[1], cb)
    e1 = (np.einsum("wij,ij->w", gu, ctx.rot_h1_a, optimize=True)
          + np.einsum("wij,ij->w", gd, ctx.rot_h1_b, optimize=True))
    fu = np.einsum("gij,wkj->wgik", ctx.rot_chol_a, gu, optimize=True)
    fd = np.einsum("gij,wkj->wgik", ctx.rot_chol_b, gd, optimize=True)
    cu = np.einsum("wgii->wg", fu, optimize=True)
    cd = np.einsum("wgii->wg", fd, optimize=True)
    exu = np.einsum("wgij,wgji->w", fu, fu, optimize=True)
    exd = np.einsum("wgij,wgji->w", fd, fd, optimize=True)
    e2 = (np.sum(cu * cu, 1) + np.sum(cd * cd, 1) + 2.0 * np.sum(cu * cd, 1) - exu - exd) / 2.0
    return h0 + e1 + e2


# ---------------------------------------------------------------------------
# propagator pieces (trot/prop/chol_afqmc_ops.py, trot/prop/utils.py)
# ---------------------------------------------------------------------------


def make_vhs(xbar: np.ndarray, chol_flat: np.ndarray, norb: int) -> np.ndarray:
    """trot/prop/chol_afqmc_ops.py:83-87: V = (Re x @ Lflat + i Im x @ Lflat).reshape(N, N).

    A float32 chol_flat is the reference's mixed mode (afqmc.py:177 stores chol_flat as float32;
    chol_afqmc_ops.py:247-257 casts x to complex64): the two products run in float32 and V is cast back
    to complex128 before the Taylor series (chol_afqmc_ops.py:160).
    """
    if chol_flat.dtype == np.float32:
        x = xbar.astype(np.complex64)
        v_re = np.ascontiguousarray(np.real(x)) @ chol_flat
        v_im = np.ascontiguousarray(np.imag(x)) @ chol_flat
        return (v_re.astype(np.float64) + 1.0j * v_im.astype(np.float64)).reshape(xbar.shape[0], norb, norb)
    # contiguous copies keep numpy on its BLAS path (a strided view of a complex array is not)
    v_re = np.ascontiguousarray(np.real(xbar)) @ chol_flat
    v_im = np.ascontiguousarray(np.imag(xbar)) @ chol_flat
    return (v_re + 1.0j * v_im).reshape(xbar.shape[0], norb, norb)


def taylor_expm_action(a: complex, mat: np.ndarray, vecs: np.ndarray, n_terms: int) -> np.ndarray:
    """trot/prop/utils.py:20-29: out = w; term <- (a/k) mat @ term; out += term, k = 1..K."""
    out = vecs
    term = vecs
    for k in range(1, n_terms + 1):
        term = (a / k) * (mat @ term)
        out = out + term
    return out


def apply_trotter_r(w: np.ndarray, xbar: np.ndarray, ctx: PropCtx, n_terms: int) -> np.ndarray:
    """trot/prop/chol_afqmc_ops.py:198-208 (restricted walkers)."""
    e = ctx.exp_h1_half
    w1 = e @ w  # chol_afqmc_ops.py:130-131
    v = make_vhs(xbar, ctx.chol_flat, ctx.norb)
    w2 = taylor_expm_action(1.0j * ctx.sqrt_dt, v, w1, n_terms)  # chol_afqmc_ops.py:152-162
    return e @ w2


def apply_trotter_u(w: tuple, xbar: np.ndarray, ctx: PropCtx, n_terms: int) -> tuple:
    """trot/prop/chol_afqmc_ops.py:211-222: one V shared by both spins."""
    e = ctx.exp_h1_half
    v = make_vhs(xbar, ctx.chol_flat, ctx.norb)
    a = 1.0j * ctx.sqrt_dt
    out = []
    for ws in w:
        w1 = e @ ws
        w2 = taylor_expm_action(a, v, w1, n_terms)
        out.append(e @ w2)
    return tuple(out)


# ---------------------------------------------------------------------------
# the phaseless step (trot/prop/afqmc.py:79-144)
# ---------------------------------------------------------------------------


def field_shifts(force_bias: np.ndarray, ctx: PropCtx) -> np.ndarray:
    """trot/prop/afqmc.py:99: fs = -sqrt(dt) (i fb - mf)."""
    return -ctx.sqrt_dt * (1.0j * force_bias - ctx.mf_shifts)


def shift_terms(fields: np.ndarray, force_bias: np.ndarray, ctx: PropCtx):
    """trot/prop/afqmc.py:99-103: returns (xbar, shift_term, fb_term)."""
    fs = field_shifts(force_bias, ctx)
    xbar = fields - fs
    shift_term = np.sum(xbar * ctx.mf_shifts, axis=1)
    fb_term = np.sum(fields * fs - 0.5 * fs * fs, axis=1)
    return xbar, shift_term, fb_term


def importance_exponent(shift_term: np.ndarray, fb_term: np.ndarray, pop_shift: float, ctx: PropCtx) -> np.ndarray:
    """trot/prop/afqmc.py:114-116: exponent = -sqrt(dt) shift_term + fb_term + dt (E_shift + h0_prop)."""
    return -ctx.sqrt_dt * shift_term + fb_term + ctx.dt * (pop_shift + ctx.h0_prop)


def phaseless_theta(shift_term: np.ndarray, ratio: np.ndarray, ctx: PropCtx) -> np.ndarray:
    """trot/prop/afqmc.py:119: theta = angle(exp(-sqrt(dt) shift_term) ratio), in (-pi, pi]."""
    return np.angle(np.exp(-ctx.sqrt_dt * shift_term) * ratio)


def phaseless_update(
    weights: np.ndarray,
    overlaps_old: np.ndarray,
    overlaps_new: np.ndarray,
    shift_term: np.ndarray,
    fb_term: np.ndarray,
    pop_shift: float,
    e_estimate: float,
    node_encounters: int,
    ctx: PropCtx,
    params: QmcParams,
):
    """trot/prop/afqmc.py:111-134.

    Returns (weights_new, node_encounters_new, pop_shift_new).  Masking order:
    non-finite or < floor -> 0; count <= 0 as nodes; > cap -> 0; multiply;
    weights > cap -> 0 (afqmc.py:122-130).
    """
    with np.errstate(all="ignore"):
        ratio = overlaps_new / overlaps_old
        exponent = importance_exponent(shift_term, fb_term, pop_shift, ctx)
        imp_fun = np.exp(exponent) * ratio
        theta = phaseless_theta(shift_term, ratio, ctx)
        imp_ph = np.abs(imp_fun) * np.cos(theta)
        imp_ph = np.where(~np.isfinite(imp_ph) | (imp_ph < params.weight_floor), 0.0, imp_ph)
        nodes = int(node_encounters) + int(np.sum(imp_ph <= 0.0))
        imp_ph = np.where(imp_ph > params.weight_cap, 0.0, imp_ph)
        weights_new = weights * imp_ph
        weights_new = np.where(weights_new > params.weight_cap, 0.0, weights_new)
        avg_w = max(float(np.mean(weights_new)), 1.0e-300)
        pop_shift_new = e_estimate - params.pop_control_damping * (np.log(avg_w) / ctx.dt)
    return weights_new, nodes, float(pop_shift_new)


@dataclass
class RhfOps:
    """Batched force-bias / overlap / energy closures for one trial kind (the MeasOps role)."""

    overlap: Callable[[Any], np.ndarray]
    force_bias: Callable[[Any], np.ndarray]
    energy: Callable[[Any], np.ndarray]
    walker_kind: str = "restricted"
    observables: dict = field(default_factory=dict)


def make_rhf_ops(h0: float, h1: np.ndarray, chol: np.ndarray, c: np.ndarray, walker_kind: str = "restricted",
                 memory_mode: str = "high") -> RhfOps:
    """trot/meas/rhf.py:322-367 (make_rhf_meas_ops) for restricted / unrestricted walkers."""
    ctx = build_meas_ctx_rhf(h1, chol, c, memory_mode)
    if walker_kind == "restricted":
        return RhfOps(
            overlap=lambda w: overlap_r(w, c),
            force_bias=lambda w: force_bias_rw(w, ctx, c),
            energy=lambda w: energy_rw(w, h0, ctx, c),
            walker_kind="restricted",
            observables={"rdm1": lambda w: rdm1_rw(w, c),
                         "density_corr": lambda w: density_corr_from_rdm1(rdm1_rw(w, c))},
        )
    if walker_kind == "unrestricted":
        return RhfOps(
            overlap=lambda w: overlap_u(w, c),
            force_bias=lambda w: force_bias_uw(w, ctx, c),
            energy=lambda w: energy_uw(w, h0, ctx, c),
            walker_kind="unrestricted",
            observables={"rdm1": lambda w: rdm1_uw(w, c),
                         "density_corr": lambda w: density_corr_from_rdm1(rdm1_uw(w, c))},
        )
    raise ValueError(f"unknown walker_kind: {walker_kind}")


def make_uhf_ops(h0: float, h1: np.ndarray, chol: np.ndarray, ca: np.ndarray, cb: np.ndarray) -> RhfOps:
    """trot/meas/uhf.py:339-379 (make_uhf_meas_ops), unrestricted walkers."""
    ctx = build_meas_ctx_uhf(h1, chol, ca, cb)
    return RhfOps(
        overlap=lambda w: overlap_u(w, ca, cb),
        force_bias=lambda w: force_bias_uhf(w, ctx, ca, cb),
        energy=lambda w: energy_uhf(w, h0, ctx, ca, cb),
        walker_kind="unrestricted",
        observables={"rdm1": lambda w: rdm1_uw(w, ca, cb),
                     "density_corr": lambda w: density_corr_from_rdm1(rdm1_uw(w, ca, cb))},
    )


def afqmc_step(state: PropState, fields: np.ndarray, ctx: PropCtx, ops: RhfOps, params: QmcParams) -> PropState:
    """trot/prop/afqmc.py:79-144 with the auxiliary fields (W, G) given instead of drawn."""
    fb = ops.force_bias(state.walkers)  # afqmc.py:95-98
    xbar, shift_term, fb_term = shift_terms(fields, fb, ctx)
    if ops.walker_kind == "unrestricted":
        w_new = apply_trotter_u(state.walkers, xbar, ctx, params.n_exp_terms)
    else:
        w_new = apply_trotter_r(state.walkers, xbar, ctx, params.n_exp_terms)
    ov_new = ops.overlap(w_new)  # afqmc.py:108-110
    weights, nodes, pop_shift = phaseless_update(
        state.weights, state.overlaps, ov_new, shift_term, fb_term,
        state.pop_control_ene_shift, state.e_estimate, state.node_encounters, ctx, params,
    )
    return PropState(
        walkers=w_new,
        weights=weights,
        overlaps=ov_new,
        pop_control_ene_shift=pop_shift,
        e_estimate=state.e_estimate,
        node_encounters=nodes,
    )


# ---------------------------------------------------------------------------
# walkers: init, QR, SR (trot/walkers.py)
# ---------------------------------------------------------------------------


def natorbs(dm: np.ndarray, n_occ: int) -> np.ndarray:
    """trot/walkers.py:14-17."""
    dm = 0.5 * (dm + dm.conj().T)
    vecs = np.linalg.eigh(dm)[1][:, ::-1]
    return vecs[:, :n_occ]


def init_walkers(rdm1: np.ndarray, nelec: tuple[int, int], n_walkers: int, walker_kind: str):
    """trot/walkers.py:20-78 (restricted / unrestricted)."""
    if walker_kind == "restricted":
        w0 = natorbs(rdm1[0] + rdm1[1], nelec[0]) + 0.0j
        return np.broadcast_to(w0, (n_walkers, *w0.shape)).copy()
    if walker_kind == "unrestricted":
        wu = natorbs(rdm1[0], nelec[0]) + 0.0j
        wd = natorbs(rdm1[1], nelec[1]) + 0.0j
        return (np.broadcast_to(wu, (n_walkers, *wu.shape)).copy(),
                np.broadcast_to(wd, (n_walkers, *wd.shape)).copy())
    raise ValueError(f"unknown walker_kind: {walker_kind}")


def qr_phase(mat: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """trot/walkers.py:137-148: reduced QR with diag(R) made real >= 0; returns (q, det_r)."""
    q, r = np.linalg.qr(mat, mode="reduced")
    d = np.diagonal(r, axis1=-2, axis2=-1)
    abs_d = np.abs(d)
    phase = d / np.where(abs_d == 0, 1.0, abs_d)
    q = q * np.conj(phase)[..., None, :]
    r = phase[..., :, None] * r
    det_r = np.prod(np.diagonal(r, axis1=-2, axis2=-1), axis=-1)
    return q, det_r


def orthonormalize(w, walker_kind: str):
    """trot/walkers.py:151-179 (norms discarded)."""
    if walker_kind == "unrestricted":
        return (qr_phase(w[0])[0], qr_phase(w[1])[0])
    return qr_phase(w)[0]


def sr_indices(weights: np.ndarray, zeta: float, n_walkers: int) -> np.ndarray:
    """trot/walkers.py:249-254: comb over cumsum(|w|), searchsorted side='left'."""
    cw = np.cumsum(np.abs(weights))
    tot = cw[-1]
    z = tot * (np.arange(n_walkers) + zeta) / n_walkers
    return np.searchsorted(cw, z, side="left")


def stochastic_reconfiguration(w, weights: np.ndarray, zeta: float, walker_kind: str):
    """trot/walkers.py:257-293: returns (walkers[idx], full(sum|w|/W))."""
    n = weights.shape[0]
    cw = np.cumsum(np.abs(weights))
    avg = cw[-1] / n
    idx = sr_indices(weights, zeta, n)
    w_new = (w[0][idx], w[1][idx]) if walker_kind == "unrestricted" else w[idx]
    return w_new, np.full((n,), avg, dtype=weights.dtype), idx


# ---------------------------------------------------------------------------
# block (trot/prop/blocks.py:195-287) and state init (trot/prop/afqmc.py:19-76)
# ---------------------------------------------------------------------------


def block(state: PropState, fields_seq, zeta: float, ctx: PropCtx, ops: RhfOps, params: QmcParams,
          observable_names: tuple = ()):
    """trot/prop/blocks.py:195-287 with pre-drawn fields (n_prop_steps, W, G) and SR zeta.

    Returns (state, {"energy": e_block, "weight": w_sum, "e_samples": ..., "sr_idx": ...,
    "observables": {name: weighted block mean}}).
    """
    for s in range(params.n_prop_steps):
        state = afqmc_step(state, fields_seq[s], ctx, ops, params)
    w_on = orthonormalize(state.walkers, ops.walker_kind)  # blocks.py:230
    state.walkers = w_on
    state.overlaps = ops.overlap(w_on)  # blocks.py:231-234
    e_samples = np.real(ops.energy(state.walkers))  # blocks.py:236-240
    thresh = np.sqrt(2.0 / params.dt)
    e_ref = state.e_estimate
    is_nan = ~np.isfinite(e_samples)
    with np.errstate(invalid="ignore"):
        e_samples = np.where(is_nan | (np.abs(e_samples - e_ref) > thresh), e_ref, e_samples)
    weights = np.where(is_nan, 0.0, state.weights)
    w_sum = np.sum(weights)
    w_sum_safe = 1.0 if w_sum == 0 else w_sum
    e_block = np.sum(weights * e_samples) / w_sum_safe
    if w_sum == 0:
        e_block = e_ref
    alpha = params.shift_ema
    state.weights = weights
    state.e_estimate = float((1.0 - alpha) * state.e_estimate + alpha * e_block)
    obs = {name: weighted_block_mean(ops.observables[name](state.walkers), weights)  # blocks.py:259-268
           for name in observable_names}
    w_sr, weights_sr, idx = stochastic_reconfiguration(state.walkers, state.weights, zeta, ops.walker_kind)
    state.walkers = w_sr
    state.weights = weights_sr
    state.overlaps = ops.overlap(w_sr)  # blocks.py:273-276
    return state, {"energy": float(e_block), "weight": float(w_sum), "e_samples": e_samples, "sr_idx": idx,
                   "observables": obs}


# ---------------------------------------------------------------------------
# free projection (trot/prop/afqmc_fp.py:18-72, chol_afqmc_ops_fp.py:49-74, walkers.py:182-226,
# blocks.py:415-490)
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class QmcParamsFp:
    """trot/prop/types.py:55-62 (with the base fields this restatement uses)."""

    dt: float = 0.05
    n_prop_steps: int = 20
    n_qr_blocks: int = 10
    n_blocks: int = 5
    ene0: float | None = None
    n_traj: int = 10
    n_walkers: int = 200


def build_prop_ctx_fp(h0: float, h1: np.ndarray, chol: np.ndarray, rdm1: np.ndarray, dt: float,
                      ene0: float) -> PropCtx:
    """trot/prop/chol_afqmc_ops_fp.py:49-74: h0_prop = -h0 - 1/2 sum mf^2 + ene0, the rest as the phaseless ctx."""
    ctx = build_prop_ctx(h0, h1, chol, rdm1, dt)
    ctx.h0_prop = complex(ctx.h0_prop + ene0)
    return ctx


def qr_norm(w, walker_kind: str) -> np.ndarray:
    """trot/walkers.py:182-205: R of the QR only; diag(R) multiplied by its phase; det_r = prod diag;
    restricted det_r^2, unrestricted det_u det_d."""
    def det_r(mat):
        r = np.linalg.qr(mat, mode="r")
        d = np.diagonal(r, axis1=-2, axis2=-1)
        abs_d = np.abs(d)
        phase = d / np.where(abs_d == 0, 1.0, abs_d)
        r = phase[..., :, None] * r
        return np.prod(np.diagonal(r, axis1=-2, axis2=-1), axis=-1)

    if walker_kind == "unrestricted":
        return det_r(w[0]) * det_r(w[1])
    d = det_r(w)
    return d * d


def orthogonalize(w, walker_kind: str):
    """trot/walkers.py:151-172: (Q with real positive diag R, norm = det_r^2 or det_u det_d)."""
    if walker_kind == "unrestricted":
        qu, du = qr_phase(w[0])
        qd, dd = qr_phase(w[1])
        return (qu, qd), du * dd
    q, d = qr_phase(w)
    return q, d * d


def multiply_constants(w, constants: np.ndarray, walker_kind: str):
    """trot/walkers.py:208-226: each walker's constant spread over its columns as c^(1/n_total)."""
    if walker_kind == "unrestricted":
        n_total = w[0].shape[-1] + w[1].shape[-1]
        c = (constants ** (1.0 / n_total)).reshape(-1, 1, 1)
        return (w[0] * c, w[1] * c)
    n_total = 2 * w.shape[-1]
    c = (constants ** (1.0 / n_total)).reshape(-1, 1, 1)
    return w * c


def afqmc_step_fp(state: PropState, fields: np.ndarray, zeta: float, ctx: PropCtx, walker_kind: str) -> PropState:
    """trot/prop/afqmc_fp.py:18-72 with the fields (W, G) and the comb offset given instead of drawn."""
    shift_term = fields @ ctx.mf_shifts  # afqmc_fp.py:35
    constants = np.exp(-ctx.sqrt_dt * shift_term + ctx.dt * ctx.h0_prop)  # :36
    if walker_kind == "unrestricted":
        w_new = apply_trotter_u(state.walkers, fields, ctx, 10)  # :38-40
    else:
        w_new = apply_trotter_r(state.walkers, fields, ctx, 10)
    w_new = multiply_constants(w_new, constants, walker_kind)  # :42
    norms = qr_norm(w_new, walker_kind)  # :51
    weights_new = state.weights * norms.real  # :52
    n = weights_new.shape[0]
    idx = np.minimum(sr_indices(weights_new, zeta, n), n - 1)  # :59 (a gather clamps)
    w_sr, weight_sr, _ = stochastic_reconfiguration(w_new, weights_new, zeta, walker_kind)  # :60
    weight_sr = weight_sr / norms[idx].real  # :61-62
    return PropState(walkers=w_sr, weights=weight_sr, overlaps=state.overlaps,
                     pop_control_ene_shift=state.pop_control_ene_shift, e_estimate=state.e_estimate,
                     node_encounters=state.node_encounters)


def block_fp(state: PropState, fields_seq, zetas, ctx: PropCtx, ops: RhfOps, params: QmcParamsFp):
    """trot/prop/blocks.py:415-490 with pre-drawn fields and comb offsets (n_prop_steps each)."""
    assert params.n_prop_steps % params.n_qr_blocks == 0
    per = params.n_prop_steps // params.n_qr_blocks
    s = 0
    for _ in range(params.n_qr_blocks):
        for _ in range(per):
            state = afqmc_step_fp(state, fields_seq[s], zetas[s], ctx, ops.walker_kind)
            s += 1
        q, norms = orthogonalize(state.walkers, ops.walker_kind)  # blocks.py:451-458
        state.weights = state.weights * norms.real
        state.walkers = q
    state.overlaps = ops.overlap(state.walkers)  # blocks.py:463-466
    e_samples = ops.energy(state.walkers)  # :467-470 (complex)
    thresh = np.sqrt(2.0 / params.dt)
    ene0 = params.ene0
    e_samples = np.where(np.abs(e_samples - ene0) > thresh, ene0, e_samples)  # :474-476
    wov = state.weights * state.overlaps
    w_sum = np.sum(wov)
    e_block = np.sum(wov * e_samples) / w_sum
    return state, {"energy": e_block, "weight": w_sum, "overlap": np.sum(state.overlaps),
                   "abs_overlap": np.sum(np.abs(state.overlaps))}


def init_prop_state(walkers, ops: RhfOps, n_walkers: int, initial_e_estimate: float | None = None) -> PropState:
    """trot/prop/afqmc.py:19-76: weights 1, overlaps, E_est = Re E_L(walker 0), shift = E_est."""
    overlaps = ops.overlap(walkers)
    if initial_e_estimate is None:
        w0 = (walkers[0][:1], walkers[1][:1]) if isinstance(walkers, tuple) else walkers[:1]
        initial_e_estimate = float(np.mean(np.real(ops.energy(w0))))
    return PropState(
        walkers=walkers,
        weights=np.ones((n_walkers,)),
        overlaps=overlaps,
        pop_control_ene_shift=float(initial_e_estimate),
        e_estimate=float(initial_e_estimate),
        node_encounters=0,
    )


# ---------------------------------------------------------------------------
# derivative-of-overlap oracle (trot/meas/auto.py:73-187)
# ---------------------------------------------------------------------------


def force_bias_auto(w: np.ndarray, chol: np.ndarray, c: np.ndarray, eps: float = 1e-6) -> np.ndarray:
    """trot/meas/auto.py:73-96: fb_g = d/dx <T|(1 + x L_g) w> / <T|w> at x=0 (central differences)."""
    ov0 = overlap_r(w, c)
    out = np.empty((w.shape[0], chol.shape[0]), dtype=complex)
    for g, lg in enumerate(chol):
        lw = lg @ w
        out[:, g] = (overlap_r(w + eps * lw, c) - overlap_r(w - eps * lw, c)) / (2 * eps) / ov0
    return out


def energy_auto(w: np.ndarray, h0: float, h1: np.ndarray, chol: np.ndarray, c: np.ndarray, eps: float = 1e-4) -> np.ndarray:
    """trot/meas/auto.py:145-187: (d/dx <T|(1+x h1_eff)w> + 1/2 sum_g d2/dx2 <T|(1+xL+x^2L^2/2)w>)/<T|w> + h0.

    Here h1_eff = h1 - 1/2 sum_g L_g L_g (the one-body part of H in Cholesky form).
    Both derivatives by central differences.
    """
    he = h1 - 0.5 * np.einsum("gik,gkj->ij", chol, chol, optimize=True)
    ov0 = overlap_r(w, c)
    hw = he @ w
    d1 = (overlap_r(w + eps * hw, c) - overlap_r(w - eps * hw, c)) / (2 * eps)
    d2 = np.zeros_like(ov0)
    for lg in chol:
        lw = lg @ w
        llw = lg @ lw
        sp = overlap_r(w + eps * lw + 0.5 * eps * eps * llw, c)
        sm = overlap_r(w - eps * lw + 0.5 * eps * eps * llw, c)
        d2 = d2 + (sp - 2.0 * ov0 + sm) / (eps * eps)
    return (d1 + 0.5 * d2) / ov0 + h0
