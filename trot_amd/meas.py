"""Half-rotated measurement context and batched GPU force-bias / energy kernels.

Mirrors trot/meas/rhf.py (make_rhf_meas_ops, build_meas_ctx, force_bias and
energy kernels for restricted and unrestricted walkers) and trot/meas/uhf.py
(UHF trial, unrestricted walkers).  Every kernel takes a batch of walkers;
a single walker (no leading axis) is promoted to a batch of one.

Layouts (device, float64): rot_chol (G, n, N) = C^T L_g as in the reference;
rot_chol_t (G, N, n) = its transpose per g, which is the force-bias operand
matching the transposed Green's function GT = w (C^T w)^-1; rot_h1t (N, n).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import torch

from . import _b3, _lib
from .core import MeasOps, System, k_energy, k_force_bias, o_density_corr, o_rdm1
from .trial import RhfTrial, UhfTrial, _batched, _coeffs, overlap, overlap_green


@dataclass(frozen=True)
class RhfMeasCfg:
    """trot/meas/rhf.py:107-109.  Both memory modes run the same streaming kernel."""

    memory_mode: str = "high"


def _validate_memory_mode(memory_mode: str) -> str:
    if memory_mode not in ("high", "low"):
        raise ValueError(f"Unsupported RHF measurement memory_mode: {memory_mode!r}")
    return memory_mode


@dataclass
class SpinCtx:
    rot_h1: torch.Tensor  # (n, N)
    rot_h1t: torch.Tensor  # (N, n)
    rot_chol: torch.Tensor  # (G, n, N)
    rot_chol_t: torch.Tensor  # (G, N, n)
    low: dict = field(default_factory=dict)  # reduced-precision copies, filled on demand

    @property
    def rot_chol_flat(self) -> torch.Tensor:
        """(G, n*N), the reference's flattened layout (trot/meas/rhf.py:318)."""
        return self.rot_chol.reshape(self.rot_chol.shape[0], -1)


@dataclass
class RhfMeasCtx:
    """trot/meas/rhf.py:282-305."""

    spin: SpinCtx
    cfg: RhfMeasCfg = RhfMeasCfg()

    @property
    def rot_h1(self):
        return self.spin.rot_h1

    @property
    def rot_chol(self):
        return self.spin.rot_chol

    @property
    def rot_chol_flat(self):
        return self.spin.rot_chol_flat


@dataclass
class UhfMeasCtx:
    """trot/meas/uhf.py:277-315."""

    spin_a: SpinCtx
    spin_b: SpinCtx

    @property
    def rot_h1_a(self):
        return self.spin_a.rot_h1

    @property
    def rot_h1_b(self):
        return self.spin_b.rot_h1

    @property
    def rot_chol_a(self):
        return self.spin_a.rot_chol

    @property
    def rot_chol_b(self):
        return self.spin_b.rot_chol


def _spin_ctx(h1: torch.Tensor, chol: torch.Tensor, c: torch.Tensor) -> SpinCtx:
    rc, rct = _lib.build_rot_chol(chol, c)
    rot_h1 = (c.T @ h1).contiguous()
    return SpinCtx(rot_h1=rot_h1, rot_h1t=rot_h1.T.contiguous(), rot_chol=rc, rot_chol_t=rct)


def _require_chol(ham_data):
    if ham_data.basis != "restricted":
        raise ValueError("RHF/UHF MeasOps assume HamChol.basis == 'restricted'.")
    if ham_data.chol is None:
        raise ValueError("build_meas_ctx needs ham_data.chol")
    h1 = ham_data.h1
    chol = ham_data.chol
    if not isinstance(h1, torch.Tensor) or h1.device.type != "cuda" or not isinstance(chol, torch.Tensor):
        ham_data = ham_data.to_device()
        h1, chol = ham_data.h1, ham_data.chol
    return h1, chol


def build_meas_ctx(ham_data, trial_data: RhfTrial, cfg: RhfMeasCfg = RhfMeasCfg()) -> RhfMeasCtx:
    """trot/meas/rhf.py:308-319 (C^T h1 and C^T L_g on the GPU)."""
    h1, chol = _require_chol(ham_data)
    return RhfMeasCtx(spin=_spin_ctx(h1, chol, trial_data.mo_coeff), cfg=cfg)


def build_uhf_meas_ctx(ham_data, trial_data: UhfTrial) -> UhfMeasCtx:
    """trot/meas/uhf.py:318-336."""
    h1, chol = _require_chol(ham_data)
    return UhfMeasCtx(spin_a=_spin_ctx(h1, chol, trial_data.mo_coeff_a),
                      spin_b=_spin_ctx(h1, chol, trial_data.mo_coeff_b))


class Engine:
    """Batched GPU evaluation shared by the kernels and the propagator.

    kind: "rr" (RHF trial, restricted walkers), "ru" (RHF trial, unrestricted
    walkers) or "uu" (UHF trial, unrestricted walkers).
    """

    observables = (o_rdm1, o_density_corr)  # names observable_from_green evaluates

    def __init__(self, kind: str, precision: str = "f64"):
        if precision not in ("f64", "f32", "bf16x3"):
            raise ValueError(f"unknown precision: {precision}")
        self.kind = kind
        self.precision = precision

    def spins(self, meas_ctx):
        if isinstance(meas_ctx, UhfMeasCtx):
            return meas_ctx.spin_a, meas_ctx.spin_b
        return meas_ctx.spin, meas_ctx.spin

    @staticmethod
    def _f32(spin: SpinCtx, name: str) -> torch.Tensor:
        """FP32 copy of a Cholesky-side operand, made once per context."""
        if name not in spin.low:
            spin.low[name] = getattr(spin, name).to(torch.float32).contiguous()
        return spin.low[name]

    @staticmethod
    def _b3(spin: SpinCtx) -> torch.Tensor:
        """Split-BF16 planes of rot_chol for the batched exchange GEMM, made once per context."""
        if "rot_chol_b3" not in spin.low:
            spin.low["rot_chol_b3"] = _lib.rot_chol_b3(spin.rot_chol.contiguous())
        return spin.low["rot_chol_b3"]

    def overlap_green(self, walkers, trial_data, out=None):
        return overlap_green(walkers, trial_data, want_green=True, out=out)

    def observable_from_green(self, name: str, greens, meas_ctx, trial_data) -> torch.Tensor:
        """Per-walker block observables from the transposed half Green's functions.

        "rdm1": (W, 2, N, N) complex, spin s = GT_s C_s^T = w_s (C_s^T w_s)^-1 C_s^T (restricted walkers
        give the same matrix for both spins); trot/meas/rhf.py:191-216, trot/meas/uhf.py:97-110.
        "density_corr": (W, 3, N, N) complex (uu, ud, dd) from the two spins' rdm1;
        trot/meas/rhf.py:237-261, trot/meas/uhf.py:129-160.
        """
        if name not in (o_rdm1, o_density_corr):
            raise KeyError(f"missing requested observable '{name}'. available: [{o_density_corr}, {o_rdm1}]")
        ca, cb = _coeffs(trial_data)
        g_a, g_b = greens
        W, N, _ = g_a.shape
        dm = torch.empty((W, 2, N, N), dtype=torch.complex128, device=g_a.device)
        _lib.rdm1_from_green(g_a, ca, dm, 0)
        _lib.rdm1_from_green(g_a if g_b is None else g_b, cb, dm, 1)
        return dm if name == o_rdm1 else _lib.density_corr(dm)

    def observable_mean_from_green(self, name: str, greens, weights, wsum, meas_ctx, trial_data):
        """The block value of a linear observable without per-walker samples, or None for others.

        rdm1 is linear in GT: sum_k w_k GT_k C^T / sum_k w_k = (sum_k w_k GT_k / sum_k w_k) C^T, so the block
        mean needs one weighted mean of GT (W N n values) and one product, not the W x 2 x N x N samples.
        """
        if name != o_rdm1:
            return None
        ca, cb = _coeffs(trial_data)
        g_a, g_b = greens
        N = g_a.shape[1]
        ma = _lib.weighted_mean(g_a, weights, wsum).unsqueeze(0).contiguous()
        mb = ma if g_b is None else _lib.weighted_mean(g_b, weights, wsum).unsqueeze(0).contiguous()
        dm = torch.empty((1, 2, N, N), dtype=torch.complex128, device=g_a.device)
        _lib.rdm1_from_green(ma, ca, dm, 0)
        _lib.rdm1_from_green(mb, cb, dm, 1)
        return dm[0]

    def force_bias_from_green(self, greens, meas_ctx, out: torch.Tensor | None = None) -> torch.Tensor:
        """(W, G) force bias from the transposed half Green's functions; ``out`` (split-BF16 form) receives it."""
        sa, sb = self.spins(meas_ctx)
        alpha = 2.0 if self.kind == "rr" else 1.0  # trot/meas/rhf.py:124-125 (both spins); 150-153; uhf.py:55-57
        if self.precision == "f64":
            fb = _lib.force_bias(greens[0], sa.rot_chol_t, alpha)
            if self.kind != "rr":
                fb = _lib.force_bias(greens[1], sb.rot_chol_t, 1.0, out=fb, accumulate=True)
            return fb
        p = self.precision
        if p == "bf16x3":
            # NT GEMM on split-bf16 planes: A = GT (W x N n, complex), B = R' (G x N n), split K
            W = greens[0].shape[0]
            G = sa.rot_chol_t.shape[0]
            fb = out if out is not None else torch.empty((W, G), dtype=torch.complex128, device=greens[0].device)
            for i, (g, s) in enumerate(((greens[0], sa), (greens[1], sb))):
                if i == 1 and self.kind == "rr":
                    break
                if "rot_chol_t_b3" not in s.low:
                    s.low["rot_chol_t_b3"] = _b3.split(s.rot_chol_t.reshape(G, -1))
                gp = _b3.split(g.reshape(W, -1))
                _b3.gemm_nt(gp, s.low["rot_chol_t_b3"], fb, alpha if i == 0 else 1.0, 0.0 if i == 0 else 1.0,
                            ksplit=8)
            return fb
        fb = _lib.force_bias_lp(greens[0], self._f32(sa, "rot_chol_t"), alpha, p)
        if self.kind != "rr":
            fb = _lib.force_bias_lp(greens[1], self._f32(sb, "rot_chol_t"), 1.0, p, out=fb, accumulate=True)
        return fb

    def energy_from_green(self, greens, h0: float, meas_ctx) -> torch.Tensor:
        sa, sb = self.spins(meas_ctx)
        if self.precision == "f64":
            if self.kind == "rr":
                return _lib.local_energy(sa.rot_chol, greens[0], sa.rot_h1t, h0)
            return _lib.local_energy(sa.rot_chol, greens[0], sa.rot_h1t, h0, sb.rot_chol, greens[1], sb.rot_h1t)
        p = self.precision
        n_max = max(g.shape[2] for g in greens if g is not None)
        if p == "bf16x3" and n_max <= _lib.MAX_NOCC:  # the batched exchange GEMM works on 64-row blocks
            G = sa.rot_chol.shape[0]
            if self.kind == "rr":
                return _lib.local_energy_b3(self._b3(sa), G, greens[0], sa.rot_h1t, h0)
            return _lib.local_energy_b3(self._b3(sa), G, greens[0], sa.rot_h1t, h0, self._b3(sb), greens[1],
                                        sb.rot_h1t)
        if self.kind == "rr":
            return _lib.local_energy_lp(self._f32(sa, "rot_chol"), greens[0], sa.rot_h1t, h0, p)
        return _lib.local_energy_lp(self._f32(sa, "rot_chol"), greens[0], sa.rot_h1t, h0, p,
                                    self._f32(sb, "rot_chol"), greens[1], sb.rot_h1t)


def _h0(ham_data) -> float:
    h0 = ham_data.h0
    if isinstance(h0, torch.Tensor):
        h0 = h0.detach().cpu().item()
    return float(complex(h0).real)


def _make_kernels(engine: Engine):
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

    return force_bias_kernel, energy_kernel


def _make_observables(engine: Engine) -> dict:
    def kernel(name):
        def obs_kernel(walker, ham_data, meas_ctx, trial_data):
            w, squeeze = _batched(walker)
            _, greens = engine.overlap_green(w, trial_data)
            out = engine.observable_from_green(name, greens, meas_ctx, trial_data)
            return out[0] if squeeze else out

        obs_kernel.__name__ = f"{name}_kernel"
        return obs_kernel

    return {o_rdm1: kernel(o_rdm1), o_density_corr: kernel(o_density_corr)}


def make_rhf_meas_ops(sys: System, memory_mode: str = "high", *, precision: str = "f64") -> MeasOps:
    """trot/meas/rhf.py:322-367 for restricted and unrestricted walkers.

    ``precision`` ("f64", "f32", "bf16x3") selects the MFMA precision of the
    force-bias and exchange GEMMs; Green's functions, determinants and all
    sums stay FP64.
    """
    cfg = RhfMeasCfg(memory_mode=_validate_memory_mode(memory_mode))
    wk = sys.walker_kind
    if wk == "restricted":
        engine = Engine("rr", precision)
    elif wk == "unrestricted":
        engine = Engine("ru", precision)
    elif wk == "generalized":
        raise NotImplementedError("generalized walkers are not part of the trot_amd GPU path")
    else:
        raise ValueError(f"unknown walker_kind: {sys.walker_kind}")
    fb, en = _make_kernels(engine)
    return MeasOps(
        overlap=overlap,
        build_meas_ctx=lambda ham_data, trial_data: build_meas_ctx(ham_data, trial_data, cfg),
        kernels={k_force_bias: fb, k_energy: en},
        observables=_make_observables(engine),
        engine=engine,
    )


def make_uhf_meas_ops(sys: System, *, precision: str = "f64") -> MeasOps:
    """trot/meas/uhf.py:339-379 (unrestricted walkers)."""
    wk = sys.walker_kind
    if wk in ("restricted", "generalized"):
        raise NotImplementedError(f"UHF trial with {wk} walkers is not part of the trot_amd GPU path")
    if wk != "unrestricted":
        raise ValueError(f"unknown walker_kind: {sys.walker_kind}")
    engine = Engine("uu", precision)
    fb, en = _make_kernels(engine)
    return MeasOps(
        overlap=overlap,
        build_meas_ctx=build_uhf_meas_ctx,
        kernels={k_force_bias: fb, k_energy: en},
        observables=_make_observables(engine),
        engine=engine,
    )
