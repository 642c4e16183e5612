"""CPU FP64 oracle for the restricted CISD trial: overlap, force bias, local energy, mixed 1-RDM.

TEST INFRASTRUCTURE ONLY (the same rules as ``oracle/afqmc_oracle.py``): only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it.

It restates, term by term in NumPy complex128 / float64, the reference's CISD trial and
measurement kernels for restricted walkers in a restricted basis:

* ``overlap``       trot/trial/cisd.py:148-166 (``_overlap_r_with_doubles``; the high, realimag and
                    low doubles contractions are the same sum, :99-145);
* ``force_bias``    trot/meas/cisd.py:548-564 (``force_bias_kernel_rw_rh_high``; make_cisd_meas_ops
                    picks it for both memory modes, :842-844) with its common terms :455-479
                    (``_force_bias_common_terms_high_realimag``);
* ``energy``        trot/meas/cisd.py:608-769 (``energy_kernel_rw_rh``, the ``memory_mode="high"``
                    branch; the low branch is the same sum taken one Cholesky vector at a time);
* ``rdm1``          trot/meas/cisd.py:772-820 (``rdm1_kernel_rw``);
* ``build_meas_ctx`` trot/meas/cisd.py:92-110.

The arithmetic is the reference's FP64 configuration (``make_cisd_meas_ops(mixed_precision=False,
testing=True)``: every ``mixed_*`` dtype float64 / complex128).  Nothing here assumes the
Cholesky vectors or h1 symmetric.

Pinned (tests/test_cisd_pins.py) by: brute-force Fock-space evaluation of <T|O|w>/<T|w> with
|T> = (1 + sum c_ia E_ai + 1/2 sum c_iajb E_ai E_bj)|HF> (oracle/fock.py) for the overlap, force
bias, energy and 1-RDM; the derivative-of-overlap definition of the force bias; the reduction to
the RHF oracle when ci1 = ci2 = 0; and the reference's own invariance tests (an active-space trial
equals its zero-padded full-space twin, tests/test_cisd.py:106-157).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class CisdTrial:
    """trot/trial/cisd.py:14-88: ci1 (nocc_act, nvir_act), ci2 (nocc_act, nvir_act, nocc_act, nvir_act);
    orbital layout [core | active occ | active vir | outer]."""

    ci1: np.ndarray
    ci2: np.ndarray
    nocc_t_core: int = 0
    nvir_t_outer: int = 0

    @property
    def nocc(self) -> int:
        return int(self.ci1.shape[0])

    @property
    def nvir(self) -> int:
        return int(self.ci1.shape[1])

    @property
    def nocc_full(self) -> int:
        return self.nocc_t_core + self.nocc

    @property
    def norb(self) -> int:
        return self.nocc_full + self.nvir + self.nvir_t_outer

    @property
    def occ_act(self) -> slice:
        return slice(self.nocc_t_core, self.nocc_full)

    @property
    def vir_act(self) -> slice:
        return slice(self.nocc_full, self.nocc_full + self.nvir)


def get_rdm1(trial: CisdTrial) -> np.ndarray:
    """trot/trial/cisd.py:91-96: the reference determinant's density, both spins."""
    dm = np.diag((np.arange(trial.norb) < trial.nocc_full).astype(float))
    return np.stack([dm, dm])


def greens(w: np.ndarray, nocc: int) -> np.ndarray:
    """trot/meas/cisd.py:19-21, batched: solve(wocc^T, w^T) -> (W, nocc, N)."""
    wocc = w[:, :nocc, :]
    return np.linalg.solve(np.swapaxes(wocc, 1, 2), np.swapaxes(w, 1, 2))


def overlap(w: np.ndarray, trial: CisdTrial) -> np.ndarray:
    """trot/trial/cisd.py:148-166 (doubles: :99-100)."""
    green = greens(w, trial.nocc_full)
    o0 = np.linalg.det(w[:, : trial.nocc_full, :]) ** 2
    x = green[:, trial.occ_act, trial.vir_act]
    o1 = np.einsum("ia,wia->w", trial.ci1, x)
    o2 = 2.0 * np.einsum("iajb,wia,wjb->w", trial.ci2, x, x, optimize=True) - np.einsum(
        "iajb,wib,wja->w", trial.ci2, x, x, optimize=True)
    return (1.0 + 2.0 * o1 + o2) * o0


def active_green_blocks(green: np.ndarray, trial: CisdTrial):
    """trot/meas/cisd.py:24-44: green_act (W, nocc_act, N), green_occ (W, nocc_act, nvir_act), greenp (W, N, nvir_act)."""
    W = green.shape[0]
    green_act = green[:, trial.occ_act, :]
    green_occ = green[:, trial.occ_act, trial.vir_act]
    greenp = np.zeros((W, trial.norb, trial.nvir), dtype=green.dtype)
    greenp[:, : trial.nocc_full, :] = green[:, :, trial.vir_act]
    greenp[:, trial.vir_act, :] = -np.eye(trial.nvir)
    return green_act, green_occ, greenp


@dataclass(frozen=True)
class CisdMeasCtx:
    """trot/meas/cisd.py:63-79."""

    rot_chol: np.ndarray  # (G, nocc_full, N)
    lci1: np.ndarray  # (G, N, nocc_act)


def build_meas_ctx(chol: np.ndarray, trial: CisdTrial) -> CisdMeasCtx:
    """trot/meas/cisd.py:92-110."""
    rot_chol = chol[:, : trial.nocc_full, :]
    lci1 = np.einsum("git,pt->gip", chol[:, :, trial.vir_act], trial.ci1, optimize=True)
    return CisdMeasCtx(rot_chol=rot_chol, lci1=lci1)


def _ci2g(ci2: np.ndarray, green_occ: np.ndarray):
    """trot/meas/cisd.py:179-198: ci2g_c[q,u] = sum ci2[p,t,q,u] G[p,t]; ci2g_e[q,t] = sum ci2[p,t,q,u] G[p,u]."""
    c = np.einsum("ptqu,wpt->wqu", ci2, green_occ, optimize=True)
    e = np.einsum("ptqu,wpu->wqt", ci2, green_occ, optimize=True)
    return c, e


def _chol_contract(chol: np.ndarray, mat: np.ndarray) -> np.ndarray:
    """trot/meas/cisd.py:124-136: sum_ij chol[g,i,j] mat[w,i,j]."""
    return np.einsum("gij,wij->wg", chol, mat, optimize=True)


def force_bias(w: np.ndarray, chol: np.ndarray, ctx: CisdMeasCtx, trial: CisdTrial) -> np.ndarray:
    """trot/meas/cisd.py:548-564 with the common terms of :455-479."""
    ci1, ci2 = trial.ci1, trial.ci2
    green = greens(w, trial.nocc_full)
    green_act, green_occ, greenp = active_green_blocks(green, trial)
    lg = np.einsum("gpj,wpj->wg", ctx.rot_chol, green, optimize=True)
    ci1g = np.einsum("pt,wpt->w", ci1, green_occ)
    ci1gp = np.einsum("pt,wit->wpi", ci1, greenp, optimize=True)
    gci1gp = np.einsum("wpj,wpi->wij", green_act, ci1gp, optimize=True)
    ci2g_c, ci2g_e = _ci2g(ci2, green_occ)
    cisd_green_c = greenp @ np.swapaxes(ci2g_c, 1, 2) @ green_act
    cisd_green_e = greenp @ np.swapaxes(ci2g_e, 1, 2) @ green_act
    cisd_green = -4.0 * cisd_green_c + 2.0 * cisd_green_e
    ci2g = 4.0 * ci2g_c - 2.0 * ci2g_e
    gci2g = np.einsum("wqu,wqu->w", ci2g, green_occ)
    ovl = 1.0 + 2.0 * ci1g + 0.5 * gci2g
    fb_0 = 2.0 * lg
    fb_1_1 = 4.0 * ci1g[:, None] * lg
    fb_2_1 = lg * gci2g[:, None]
    fb_corr = _chol_contract(chol, cisd_green - 2.0 * gci1gp)
    return (fb_0 + fb_1_1 + fb_2_1 + fb_corr) / ovl[:, None]


def energy(w: np.ndarray, h0: float, h1: np.ndarray, chol: np.ndarray, ctx: CisdMeasCtx,
           trial: CisdTrial) -> np.ndarray:
    """trot/meas/cisd.py:608-769, memory_mode "high" (:695-700, :702-718, :745-755)."""
    ci1, ci2 = trial.ci1, trial.ci2
    act = trial.occ_act
    green = greens(w, trial.nocc_full)
    green_act, green_occ, greenp = active_green_blocks(green, trial)
    T = lambda a: np.swapaxes(a, 1, 2)  # noqa: E731
    rot_chol = ctx.rot_chol

    e0 = h0
    # one body (:624-649)
    hg = np.einsum("pj,wpj->w", h1[: trial.nocc_full, :], green)
    e1_0 = 2.0 * hg
    ci1g = np.einsum("pt,wpt->w", ci1, green_occ)
    e1_1_1 = 4.0 * ci1g * hg
    gpci1 = greenp @ ci1.T
    ci1_green = gpci1 @ green_act
    e1_1_2 = -2.0 * np.einsum("ij,wij->w", h1, ci1_green)
    e1_1 = e1_1_1 + e1_1_2
    ci2g_c, ci2g_e = _ci2g(ci2, green_occ)
    ci2_green_c = (greenp @ T(ci2g_c)) @ green_act
    ci2_green_e = (greenp @ T(ci2g_e)) @ green_act
    ci2_green = 2.0 * ci2_green_c - 1.0 * ci2_green_e
    ci2g = 2.0 * ci2g_c - 1.0 * ci2g_e
    gci2g = np.einsum("wqu,wqu->w", ci2g, green_occ)
    e1_2_1 = 2.0 * hg * gci2g
    e1_2_2 = -2.0 * np.einsum("ij,wij->w", h1, ci2_green)
    e1_2 = e1_2_1 + e1_2_2
    e1 = e1_0 + e1_1 + e1_2

    # two body (:651-767)
    lg = np.einsum("gpj,wpj->wg", rot_chol, green, optimize=True)
    e2_0_1 = 2.0 * np.einsum("wg,wg->w", lg, lg)
    lg1 = np.einsum("gpj,wqj->wgpq", rot_chol, green, optimize=True)
    e2_0_2 = -np.einsum("wgpq,wgqp->w", lg1, lg1, optimize=True)
    e2_0 = e2_0_1 + e2_0_2

    e2_1_1 = 2.0 * e2_0 * ci1g
    lci1g = _chol_contract(chol, ci1_green)
    e2_1_2 = -2.0 * np.einsum("wg,wg->w", lci1g, lg)
    ci1g1 = np.einsum("at,wpt->wap", ci1, green[:, :, trial.vir_act])
    e2_1_3_1 = np.einsum("wgpq,wgqa,wap->w", lg1, lg1[:, :, :, act], ci1g1, optimize=True)
    lci1g_mat = np.einsum("gia,wqi->wgaq", ctx.lci1, green, optimize=True)
    e2_1_3_2 = -np.einsum("wgaq,wgqa->w", lci1g_mat, lg1[:, :, :, act], optimize=True)
    e2_1_3 = e2_1_3_1 + e2_1_3_2
    e2_1 = e2_1_1 + 2.0 * (e2_1_2 + e2_1_3)

    e2_2_1 = e2_0 * gci2g
    lci2g = _chol_contract(chol, ci2_green)
    e2_2_2_1 = -np.einsum("wg,wg->w", lci2g, lg)
    lci2_green = np.einsum("gpi,wji->wgpj", rot_chol, ci2_green, optimize=True)
    gl = np.einsum("wpj,gji->wgpi", green, chol, optimize=True)  # :401-411
    e2_2_2_2 = 0.5 * np.einsum("wgpi,wgpi->w", gl, lci2_green, optimize=True)
    glgp = np.einsum("wgpi,wit->wgpt", gl, greenp, optimize=True)[:, :, act, :]
    l2ci2_1 = np.einsum("wgpt,wgqu,ptqu->w", glgp, glgp, ci2, optimize=True)  # :312-330
    l2ci2_2 = np.einsum("wgpu,wgqt,ptqu->w", glgp, glgp, ci2, optimize=True)
    e2_2_3 = 2.0 * l2ci2_1 - l2ci2_2
    e2_2_2 = 4.0 * (e2_2_2_1 + e2_2_2_2)
    e2_2 = e2_2_1 + e2_2_2 + e2_2_3
    e2 = e2_0 + e2_1 + e2_2

    ovl = 1.0 + 2.0 * ci1g + gci2g
    return (e1 + e2) / ovl + e0


def rdm1(w: np.ndarray, trial: CisdTrial) -> np.ndarray:
    """trot/meas/cisd.py:772-820: gamma = G0 - (ci1_green + ci2_green) / overlap, shape (W, 2, N, N).

    As in the reference, the CI correction enters untransposed: the exact mixed estimator
    <T|a+_p a_q|w>/<T|w> at [q, p] is G0 - (ci1_green + ci2_green)^T / overlap.  The two agree on
    every symmetric one-body operator."""
    ci1, ci2 = trial.ci1, trial.ci2
    green = greens(w, trial.nocc_full)
    green_act, green_occ, greenp = active_green_blocks(green, trial)
    W, N = w.shape[0], trial.norb
    g0 = np.zeros((W, N, N), dtype=green.dtype)
    g0[:, :, : trial.nocc_full] = np.swapaxes(green, 1, 2)
    ci1g = np.einsum("pt,wpt->w", ci1, green_occ)
    ci1_green = (greenp @ ci1.T) @ green_act
    ci2g_c, ci2g_e = _ci2g(ci2, green_occ)
    T = lambda a: np.swapaxes(a, 1, 2)  # noqa: E731
    ci2_green = 2.0 * (greenp @ T(ci2g_c)) @ green_act - (greenp @ T(ci2g_e)) @ green_act
    ci2g = 2.0 * ci2g_c - ci2g_e
    gci2g = np.einsum("wqu,wqu->w", ci2g, green_occ)
    ovl = 1.0 + 2.0 * ci1g + gci2g
    gamma = g0 - (ci1_green + ci2_green) / ovl[:, None, None]
    return np.stack([gamma, gamma], axis=1)


def slice_trial_level(trial: CisdTrial, nvir_keep: int | None) -> CisdTrial:
    """trot/trial/cisd.py:216-233."""
    if nvir_keep is None:
        return trial
    return CisdTrial(trial.ci1[:, :nvir_keep], trial.ci2[:, :nvir_keep, :, :nvir_keep], trial.nocc_t_core,
                     trial.nvir_t_outer + (trial.nvir - nvir_keep))


def force_bias_auto(w: np.ndarray, chol: np.ndarray, ovl, eps: float = 1e-5) -> np.ndarray:
    """fb_g = d/dx ln <T|(1 + x L_g) w> at x = 0 (trot/meas/auto.py:73-96), fourth-order central differences."""
    ov0 = ovl(w)
    out = np.empty((w.shape[0], chol.shape[0]), dtype=complex)
    for g, lg in enumerate(chol):
        lw = lg @ w
        d = (8.0 * (ovl(w + eps * lw) - ovl(w - eps * lw)) - (ovl(w + 2 * eps * lw) - ovl(w - 2 * eps * lw))) / (
            12.0 * eps)
        out[:, g] = d / ov0
    return out


def make_ops(h0: float, h1: np.ndarray, chol: np.ndarray, trial: CisdTrial):
    """The MeasOps role for the CISD trial (trot/meas/cisd.py:823-853), as oracle.afqmc_oracle.RhfOps."""
    from oracle.afqmc_oracle import RhfOps

    ctx = build_meas_ctx(chol, trial)
    return RhfOps(overlap=lambda w: overlap(w, trial), force_bias=lambda w: force_bias(w, chol, ctx, trial),
                  energy=lambda w: energy(w, h0, h1, chol, ctx, trial), walker_kind="restricted",
                  observables={"rdm1": lambda w: rdm1(w, trial)})
