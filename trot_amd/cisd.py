"""CISD trial for restricted walkers (trot/trial/cisd.py, trot/meas/cisd.py) on the GPU.

The reference determinant occupies the first ``nocc_full`` orbitals; the amplitudes ``ci1`` (nocc_act,
nvir_act) and ``ci2`` (nocc_act, nvir_act, nocc_act, nvir_act) refer to the active blocks of the layout
[core | active occ | active vir | outer].  Every quantity goes through the C-ABI entry points
``trot_cisd_*`` (trot_amd/csrc/cisd.hip): the half Green's function comes from the RHF kernel with the
trial I[:, :nocc_full]; both doubles contractions of the reference are one FP64 GEMM with the matrix
D[(p,t),(q,u)] = 2 ci2[p,t,q,u] - ci2[p,u,q,t], built once per trial.

Precision (``make_cisd_meas_ops``), mirroring the reference's two switches:

* ``mixed_precision=True`` (the reference's default, whose Cholesky contractions then run in float32)
  computes the force-bias GEMM on split-BF16 MFMA (16-bit significands per operand, about 1e-5
  relative, FP32 accumulation); ``mixed_precision=False`` keeps it FP64;
* ``testing=False`` (the reference's default, whose doubles term of the local energy then runs in
  float32, trot/meas/cisd.py:745-768) computes that term, sum_g x_g^T D x_g, on split-BF16 MFMA;
  ``testing=True`` keeps it FP64.

Overlaps, Green's functions and every other term of the local energy are FP64 in all modes.

Amplitudes, h1 and the Cholesky vectors must be real and the latter two symmetric (real orbitals, as
every Hamiltonian the reference builds).  Complex amplitudes raise ValueError.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Literal

import numpy as np
import torch

from . import _b3, _lib
from .core import MeasOps, System, TrialOps, k_energy, k_force_bias, o_rdm1
from .trial import _batched

_CISD_MEAS_CFG_ATTR = "_cisd_meas_cfg"


def _real(a, name: str) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        if a.is_complex():
            if torch.count_nonzero(a.imag).item():
                raise ValueError(f"trot_amd supports real CISD amplitudes only ({name})")
            a = a.real
        return a.to(device="cuda", dtype=torch.float64).contiguous()
    a = np.asarray(a)
    if np.iscomplexobj(a):
        if np.any(a.imag != 0):
            raise ValueError(f"trot_amd supports real CISD amplitudes only ({name})")
        a = a.real
    return torch.as_tensor(a, dtype=torch.float64, device="cuda").contiguous()


@dataclass(frozen=True)
class CisdTrial:
    """trot/trial/cisd.py:14-88.  ``cache`` holds D and the reference-determinant orbitals once built."""

    ci1: torch.Tensor
    ci2: torch.Tensor
    nocc_t_core: int = 0
    nvir_t_outer: int = 0
    cache: dict = field(default_factory=dict, compare=False, repr=False)

    def __post_init__(self):
        ci1, ci2 = _real(self.ci1, "ci1"), _real(self.ci2, "ci2")
        na, va = int(ci1.shape[0]), int(ci1.shape[1])
        if tuple(ci2.shape) != (na, va, na, va):
            raise ValueError(f"ci2 shape {tuple(ci2.shape)} does not match ci1 {tuple(ci1.shape)}")
        object.__setattr__(self, "ci1", ci1)
        object.__setattr__(self, "ci2", ci2)

    @property
    def nocc(self) -> int:
        """Number of active occupied orbitals."""
        return int(self.ci1.shape[0])

    @property
    def nvir(self) -> int:
        """Number of active virtual orbitals."""
        return int(self.ci1.shape[1])

    @property
    def nocc_full(self) -> int:
        return int(self.nocc_t_core + self.nocc)

    @property
    def nvir_full(self) -> int:
        return int(self.nvir + self.nvir_t_outer)

    @property
    def norb(self) -> int:
        return int(self.nocc_full + self.nvir_full)

    @property
    def occ_act_slice(self) -> slice:
        return slice(self.nocc_t_core, self.nocc_full)

    @property
    def vir_act_slice(self) -> slice:
        return slice(self.nocc_full, self.nocc_full + self.nvir)

    @property
    def norb_act(self) -> int:
        return int(self.nocc + self.nvir)

    def dmat(self) -> torch.Tensor:
        """D[(p,t),(q,u)] = 2 ci2[p,t,q,u] - ci2[p,u,q,t] (trot/trial/cisd.py:99-145 in one matrix)."""
        if "dmat" not in self.cache:
            if self.nocc < 1 or self.nvir < 1:
                raise ValueError("the CISD kernels need at least one active occupied and one active virtual orbital")
            self.cache["dmat"] = _lib.cisd_doubles_matrix(self.ci2)
        return self.cache["dmat"]

    def ref_orbitals(self) -> torch.Tensor:
        """I[:, :nocc_full]: the reference determinant as a real trial for the RHF kernels."""
        if "ref" not in self.cache:
            self.cache["ref"] = torch.eye(self.norb, dtype=torch.float64, device=self.ci1.device)[
                :, : self.nocc_full].contiguous()
        return self.cache["ref"]


def get_rdm1(trial_data: CisdTrial) -> torch.Tensor:
    """trot/trial/cisd.py:91-96."""
    occ = (torch.arange(trial_data.norb, device=trial_data.ci1.device) < trial_data.nocc_full).to(torch.float64)
    dm = torch.diag(occ)
    return torch.stack([dm, dm], dim=0)


def _overlap_parts(walkers: torch.Tensor, trial: CisdTrial, want_overlap: bool = True):
    """(overlap or None, (GT, Bs, Omega)) for a batch of restricted walkers."""
    det, gt = _lib.overlap_green(trial.ref_orbitals(), walkers, True)
    omega, ovl, bs = _lib.cisd_overlap(gt, det if want_overlap else None, trial.ci1, trial.dmat(), trial.nocc_t_core)
    return ovl, (gt, bs, omega)


def overlap_r(walker, trial_data: CisdTrial) -> torch.Tensor:
    """trot/trial/cisd.py:148-166 for one walker (N, nocc) or a batch (W, N, nocc)."""
    w, squeeze = _batched(walker)
    ov, _ = _overlap_parts(w, trial_data)
    return ov[0] if squeeze else ov


# the reference's memory / dtype variants compute the same sum (trot/trial/cisd.py:169-186)
overlap_r_high = overlap_r_high_complex = overlap_r_high_realimag = overlap_r_low = overlap_r


def make_cisd_trial_ops(sys: System, memory_mode: Literal["high", "low"] = "high") -> TrialOps:
    """trot/trial/cisd.py:189-200."""
    if sys.nup != sys.ndn:
        raise ValueError("Restricted CISD trial requires nup == ndn.")
    if sys.walker_kind.lower() != "restricted":
        raise ValueError(f"CISD trial currently supports only restricted walkers, got: {sys.walker_kind}")
    if memory_mode not in ("high", "low"):
        raise ValueError(f"Unsupported CISD trial memory_mode: {memory_mode!r}")
    return TrialOps(overlap=overlap_r, get_rdm1=get_rdm1)


def make_cisd_trial_data(data: dict, sys: System) -> CisdTrial:
    """trot/trial/cisd.py:203-213."""
    return CisdTrial(ci1=data["ci1"], ci2=data["ci2"], nocc_t_core=int(data.get("nocc_t_core", 0)),
                     nvir_t_outer=int(data.get("nvir_t_outer", 0)))


def slice_trial_level(trial: CisdTrial, nvir_keep: int | None) -> CisdTrial:
    """trot/trial/cisd.py:216-233: keep the first nvir_keep virtual amplitudes, same orbital space."""
    if nvir_keep is None:
        return trial
    return CisdTrial(ci1=trial.ci1[:, :nvir_keep], ci2=trial.ci2[:, :nvir_keep, :, :nvir_keep],
                     nocc_t_core=trial.nocc_t_core, nvir_t_outer=trial.nvir_t_outer + (trial.nvir - nvir_keep))


def slice_trial_orbital_prefix(trial: CisdTrial, nvir_keep: int | None) -> CisdTrial:
    """trot/trial/cisd.py:236-253: truncate the orbital space to nocc_full + nvir_keep orbitals."""
    if nvir_keep is None:
        return trial
    if nvir_keep < 0 or nvir_keep > trial.nvir_full:
        raise ValueError(f"nvir_keep={nvir_keep} must be in [0, {trial.nvir_full}] for this trial.")
    amp = min(int(nvir_keep), trial.nvir)
    return CisdTrial(ci1=trial.ci1[:, :amp], ci2=trial.ci2[:, :amp, :, :amp], nocc_t_core=trial.nocc_t_core,
                     nvir_t_outer=int(nvir_keep) - amp)


# ---------------------------------------------------------------------------
# measurement (trot/meas/cisd.py)
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class CisdMeasCfg:
    """trot/meas/cisd.py:47-53 (dtypes as torch dtypes)."""

    memory_mode: str = "low"
    mixed_real_dtype: torch.dtype = torch.float64
    mixed_complex_dtype: torch.dtype = torch.complex128
    mixed_real_dtype_testing: torch.dtype = torch.float32
    mixed_complex_dtype_testing: torch.dtype = torch.complex64

    @property
    def force_bias_precision(self) -> str:
        return "f64" if self.mixed_real_dtype == torch.float64 else "bf16x3"

    @property
    def doubles_precision(self) -> str:
        return "f64" if self.mixed_real_dtype_testing == torch.float64 else "bf16x3"


def get_cisd_meas_cfg(meas_ops: MeasOps) -> CisdMeasCfg | None:
    """trot/meas/cisd.py:56-60."""
    cfg = getattr(meas_ops, _CISD_MEAS_CFG_ATTR, None)
    return cfg if isinstance(cfg, CisdMeasCfg) else None


@dataclass
class CisdMeasCtx:
    """trot/meas/cisd.py:63-79, plus the device operands of the kernels.

    rot_chol (G, nocc_full, N) is a view of the Cholesky vectors; lci1 (G, N, nocc_act) is formed on first
    use only (the kernels do not read it).  lpack (G, N (N+1) / 2): upper triangles of the Cholesky
    vectors (the force-bias K index); lx (G+1, N, N): the Cholesky vectors followed by h1 (one GEMM gives
    L_g GT for every g and h1 GT)."""

    chol: torch.Tensor
    h1: torch.Tensor
    trial: CisdTrial
    cfg: CisdMeasCfg
    lpack: torch.Tensor
    lx: torch.Tensor
    low: dict = field(default_factory=dict)

    @property
    def rot_chol(self) -> torch.Tensor:
        return self.chol[:, : self.trial.nocc_full, :]

    @property
    def lci1(self) -> torch.Tensor:
        if "lci1" not in self.low:
            self.low["lci1"] = torch.einsum("git,pt->gip", self.chol[:, :, self.trial.vir_act_slice], self.trial.ci1)
        return self.low["lci1"]


def build_meas_ctx(ham_data, trial_data: CisdTrial, cfg: CisdMeasCfg = CisdMeasCfg()) -> CisdMeasCtx:
    """trot/meas/cisd.py:92-110."""
    if ham_data.basis != "restricted":
        raise ValueError("CISD MeasOps currently assumes HamChol.basis == 'restricted'.")
    if ham_data.chol is None:
        raise ValueError("the CISD measurement needs the Cholesky vectors (ham_data.chol is None)")
    chol = _real(ham_data.chol, "chol")
    h1 = _real(ham_data.h1, "h1")
    N = trial_data.norb
    if tuple(chol.shape[1:]) != (N, N) or tuple(h1.shape) != (N, N):
        raise ValueError(f"Hamiltonian with {tuple(h1.shape)} orbitals does not match the trial's {N}")
    _require_symmetric(chol, h1)
    lx = torch.cat([chol, h1[None]], dim=0).contiguous()
    return CisdMeasCtx(chol=chol, h1=h1, trial=trial_data, cfg=cfg, lpack=_lib.pack_sym(chol), lx=lx)


def _require_symmetric(chol: torch.Tensor, h1: torch.Tensor) -> None:
    """The kernels use L_g = L_g^T and h1 = h1^T (real orbitals); anything else is refused, not misread.
    Within 1e-12 of the largest element (a rotation to the MO basis leaves rounding-level asymmetry)."""
    asym = float((h1 - h1.T).abs().max())
    big = float(h1.abs().max())
    for g0 in range(0, chol.shape[0], 256):  # in slices: no (G, N, N) temporary
        c = chol[g0:g0 + 256]
        asym = max(asym, float((c - c.transpose(1, 2)).abs().max()))
        big = max(big, float(c.abs().max()))
    if asym > 1e-12 * big:
        raise ValueError("the CISD measurement needs symmetric h1 and Cholesky vectors (real orbitals)")


def slice_meas_ctx_chol(ctx: CisdMeasCtx, nchol_keep: int | None) -> CisdMeasCtx:
    """trot/meas/cisd.py:82-89."""
    if nchol_keep is None:
        return ctx
    chol = ctx.chol[:nchol_keep].contiguous()
    return CisdMeasCtx(chol=chol, h1=ctx.h1, trial=ctx.trial, cfg=ctx.cfg, lpack=ctx.lpack[:nchol_keep].contiguous(),
                       lx=torch.cat([chol, ctx.h1[None]], dim=0).contiguous())


class CisdEngine:
    """Batched GPU evaluation for the CISD trial; the cached 'green' of a state is (GT, Bs, Omega)."""

    kind = "cisd"

    def __init__(self, precision: str = "f64", doubles_precision: str = "f64"):
        if precision not in ("f64", "bf16x3") or doubles_precision not in ("f64", "bf16x3"):
            raise ValueError(f"unknown precision: {precision}, {doubles_precision}")
        self.precision = precision
        self.doubles_precision = doubles_precision

    def overlap_green(self, walkers, trial_data: CisdTrial, out=None):
        """``out`` is accepted for the engine interface; the CISD parts are allocated here."""
        if isinstance(walkers, tuple):
            raise ValueError("CISD trial supports restricted walkers only")
        return _overlap_parts(walkers, trial_data)

    @staticmethod
    def _ps(greens, ctx: CisdMeasCtx):
        gt, bs, _ = greens
        return _lib.cisd_ps(gt, bs, ctx.trial.nocc_t_core)

    def force_bias_from_green(self, greens, meas_ctx: CisdMeasCtx) -> torch.Tensor:
        gt = greens[0]
        gp = _lib.cisd_gamma(gt, self._ps(greens, meas_ctx), meas_ctx.trial.nocc_t_core, packed=True)
        if self.precision == "f64":
            return _lib.force_bias_flat(gp, meas_ctx.lpack, 2.0)
        W, G = gt.shape[0], meas_ctx.lpack.shape[0]
        if "lpack_b3" not in meas_ctx.low:
            meas_ctx.low["lpack_b3"] = _b3.split(meas_ctx.lpack)
        fb = torch.empty((W, G), dtype=torch.complex128, device=gt.device)
        _b3.gemm_nt(_b3.split(gp), meas_ctx.low["lpack_b3"], fb, 2.0, 0.0, ksplit=8)
        return fb

    def energy_from_green(self, greens, h0: float, meas_ctx: CisdMeasCtx) -> torch.Tensor:
        gt, bs, omega = greens
        t = meas_ctx.trial
        planes = None
        if self.doubles_precision == "bf16x3":
            if "dmat_b3" not in t.cache:
                t.cache["dmat_b3"] = _b3.split(t.dmat(), transposed=True)  # D^T: rows j, K index k
            planes = t.cache["dmat_b3"]
        return _lib.cisd_energy(gt, bs, self._ps(greens, meas_ctx), omega, meas_ctx.lx, t.dmat(), t.nocc_t_core, h0,
                                d_planes=planes)

    def rdm1_from_green(self, greens, meas_ctx: CisdMeasCtx) -> torch.Tensor:
        gam = _lib.cisd_gamma(greens[0], self._ps(greens, meas_ctx), meas_ctx.trial.nocc_t_core, packed=False)
        return torch.stack([gam, gam], dim=1)

    observables = (o_rdm1,)  # names observable_from_green evaluates

    def observable_from_green(self, name: str, greens, meas_ctx: CisdMeasCtx, trial_data=None) -> torch.Tensor:
        """The block observables of the CISD trial (trot/meas/cisd.py:850): "rdm1" only."""
        if name != o_rdm1:
            raise KeyError(f"missing requested observable '{name}'. available: [{o_rdm1}]")
        return self.rdm1_from_green(greens, meas_ctx)


def _h0(ham_data) -> float:
    h0 = ham_data.h0
    if isinstance(h0, torch.Tensor):
        h0 = h0.detach().cpu().item()
    return float(complex(h0).real)


def _kernels(engine: CisdEngine):
    def force_bias_kernel(walker, ham_data, meas_ctx, trial_data):
        w, squeeze = _batched(walker)
        _, greens = engine.overlap_green(w, trial_data)
        fb = engine.force_bias_from_green(greens, meas_ctx)
        return fb[0] if squeeze else fb

    def energy_kernel(walker, ham_data, meas_ctx, trial_data):
        w, squeeze = _batched(walker)
        _, greens = engine.overlap_green(w, trial_data)
        e = engine.energy_from_green(greens, _h0(ham_data), meas_ctx)
        return e[0] if squeeze else e

    def rdm1_kernel(walker, ham_data, meas_ctx, trial_data):
        w, squeeze = _batched(walker)
        _, greens = engine.overlap_green(w, trial_data)
        dm = engine.rdm1_from_green(greens, meas_ctx)
        return dm[0] if squeeze else dm

    return force_bias_kernel, energy_kernel, rdm1_kernel


_F64_ENGINE = CisdEngine("f64")
# the reference's force-bias variants (trot/meas/cisd.py:532-605) are the same sum; in FP64 here
force_bias_kernel_rw_rh, energy_kernel_rw_rh, rdm1_kernel_rw = _kernels(_F64_ENGINE)
force_bias_kernel_rw_rh_high = force_bias_kernel_rw_rh_high_complex = force_bias_kernel_rw_rh
force_bias_kernel_rw_rh_high_realimag = force_bias_kernel_rw_rh_low = force_bias_kernel_rw_rh


def make_cisd_meas_ops(sys: System, memory_mode: str = "high", mixed_precision: bool = True,
                       testing: bool = False) -> MeasOps:
    """trot/meas/cisd.py:823-853."""
    if sys.walker_kind.lower() != "restricted":
        raise ValueError(f"CISD MeasOps currently supports only restricted walkers, got: {sys.walker_kind}")
    cfg = CisdMeasCfg(
        memory_mode=memory_mode,
        mixed_real_dtype=torch.float32 if mixed_precision else torch.float64,
        mixed_complex_dtype=torch.complex64 if mixed_precision else torch.complex128,
        mixed_real_dtype_testing=torch.float64 if testing else torch.float32,
        mixed_complex_dtype_testing=torch.complex128 if testing else torch.complex64,
    )
    engine = CisdEngine(cfg.force_bias_precision, cfg.doubles_precision)
    fb, en, dm = _kernels(engine)
    ops = MeasOps(
        overlap=overlap_r,
        build_meas_ctx=lambda ham_data, trial_data: build_meas_ctx(ham_data, trial_data, cfg),
        kernels={k_force_bias: fb, k_energy: en},
        observables={o_rdm1: dm},
        engine=engine,
    )
    object.__setattr__(ops, _CISD_MEAS_CFG_ATTR, cfg)
    return ops


__all__ = [
    "CisdTrial", "get_rdm1", "overlap_r", "overlap_r_high", "overlap_r_high_complex", "overlap_r_high_realimag",
    "overlap_r_low", "make_cisd_trial_ops", "make_cisd_trial_data", "slice_trial_level",
    "slice_trial_orbital_prefix", "CisdMeasCfg", "CisdMeasCtx", "get_cisd_meas_cfg", "build_meas_ctx",
    "slice_meas_ctx_chol", "CisdEngine", "force_bias_kernel_rw_rh", "force_bias_kernel_rw_rh_high",
    "force_bias_kernel_rw_rh_high_complex", "force_bias_kernel_rw_rh_high_realimag", "force_bias_kernel_rw_rh_low",
    "energy_kernel_rw_rh", "rdm1_kernel_rw", "make_cisd_meas_ops",
]
