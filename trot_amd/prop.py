"""Phaseless AFQMC propagation on the GPU: a drop-in for trot/prop/afqmc.py.

``make_prop_ops(ham_basis, walker_kind, mixed_precision=False)`` returns
``PropOps(init_prop_state, build_prop_ctx, step)`` with the reference's
signatures (trot/prop/types.py:65-103).  Differences, all documented:

* auxiliary fields come from a field provider on the propagation context
  instead of JAX's threefry stream.  ``PropState.rng_key`` is a
  ``(seed, counter)`` pair; each step consumes one counter value.  The
  default provider draws N(0,1) on the GPU (torch Philox generator);
  ``NumpyFields`` draws the same values as ``trot_amd.synthetic.fields``
  so runs can be compared with a CPU restatement;
* the step returns a new state and never writes into the input state, so
  the reference's functional semantics hold;
* ``PropState.green`` caches the half Green's functions of ``walkers``
  (computed together with the overlaps); the next force bias reuses them
  instead of recomputing C^T w and its inverse from the same walkers.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable, NamedTuple

import numpy as np
import scipy.linalg
import torch

from . import _b3, _lib
from .core import MeasOps, System, TrialOps
from . import walkers as wk


def _random_seed() -> int:
    return int(np.random.randint(0, int(1e6)))


@dataclass(frozen=True)
class QmcParamsBase:
    """trot/prop/types.py:30-38."""

    dt: float = 0.005
    n_chunks: int = 1
    n_exp_terms: int = 6
    n_prop_steps: int = 50
    n_blocks: int = 200
    n_walkers: int = 200
    seed: int = field(default_factory=_random_seed)


@dataclass(frozen=True)
class QmcParams(QmcParamsBase):
    """trot/prop/types.py:41-47."""

    pop_control_damping: float = 0.1
    weight_floor: float = 1.0e-3
    weight_cap: float = 100.0
    shift_ema: float = 0.1
    n_eql_blocks: int = 20


class PropState(NamedTuple):
    """trot/prop/types.py:20-27, plus the Green's-function cache ``green``."""

    walkers: Any
    weights: torch.Tensor
    overlaps: torch.Tensor
    rng_key: Any
    pop_control_ene_shift: torch.Tensor
    e_estimate: torch.Tensor
    node_encounters: torch.Tensor
    green: Any = None


# ---------------------------------------------------------------------------
# auxiliary-field providers
# ---------------------------------------------------------------------------


class DeviceFields:
    """N(0,1) fields drawn on the GPU from the counter-based Philox4x64-10 stream: draw ``ctr`` of ``seed``
    (trot_philox_normal), and the SR comb offset from the same stream (trot_philox_uniform, returned as a
    one-element GPU tensor, so the block never waits on the host)."""

    def __call__(self, key, n_walkers: int, n_fields: int, device):
        seed, ctr = key
        x = _lib.philox_normal((n_walkers, n_fields), seed=int(seed), counter=int(ctr), device=device)
        return x, (seed, ctr + 1)

    def uniform(self, key):
        seed, ctr = key
        return _lib.philox_uniform(seed=int(seed), counter=int(ctr)), (seed, ctr + 1)


class StateFields:
    """The DeviceFields stream with (seed, counter) kept in a device array: each draw reads the counter
    there and advances it on the device, so a captured HIP graph draws the next numbers on every replay.
    The host key advances alongside (one per draw), as with DeviceFields."""

    def __init__(self, state: torch.Tensor):
        self.state = state

    def __call__(self, key, n_walkers: int, n_fields: int, device):
        x = _lib.philox_normal((n_walkers, n_fields), state=self.state, device=device)
        _lib.rng_advance(self.state, 1)
        return x, (key[0], key[1] + 1)

    def uniform(self, key):
        z = _lib.philox_uniform(state=self.state)
        _lib.rng_advance(self.state, 1)
        return z, (key[0], key[1] + 1)


class NumpyFields:
    """Fields drawn on the host with numpy PCG64 as trot_amd.synthetic.fields(seed, counter, W, G)."""

    def __call__(self, key, n_walkers: int, n_fields: int, device):
        from .synthetic import fields

        seed, ctr = key
        x = torch.from_numpy(fields(int(seed), int(ctr), n_walkers, n_fields)).to(device)
        return x, (seed, ctr + 1)

    def uniform(self, key):
        seed, ctr = key
        rng = np.random.default_rng([int(seed), 104729, int(ctr)])
        return float(rng.uniform()), (seed, ctr + 1)


# ---------------------------------------------------------------------------
# propagation context
# ---------------------------------------------------------------------------


@dataclass
class CholAfqmcCtx:
    """trot/prop/chol_afqmc_ops.py:16-50 plus the field provider and precision."""

    dt: float
    sqrt_dt: float
    exp_h1_half: torch.Tensor  # (N, N) float64
    mf_shifts: torch.Tensor  # (G,) complex128
    h0_prop: complex
    chol_flat: torch.Tensor  # (G, N*N) float64
    norb: int
    fields: Callable = field(default_factory=DeviceFields)
    precision: str = "f64"  # VHS GEMM precision
    taylor_precision: str = "f64"
    walker_chunk: int = 0
    taylor_hi_from: int = 0  # split-BF16 propagation: first Taylor term as one bf16 pass (0: none)
    low: dict = field(default_factory=dict)  # FP32 operand copies, kernel timer


def _total_dm(rdm1) -> torch.Tensor:
    dm = rdm1 if isinstance(rdm1, torch.Tensor) else torch.as_tensor(np.asarray(rdm1))
    if dm.dim() == 3 and dm.shape[0] == 2:
        dm = dm[0] + dm[1]  # trot/prop/chol_afqmc_ops.py:57-60
    if dm.is_complex():
        if torch.count_nonzero(dm.imag).item():
            raise ValueError("trot_amd supports real trial density matrices only")
        dm = dm.real
    return dm.to(device="cuda", dtype=torch.float64).contiguous()


def build_prop_ctx(ham_data, rdm1, dt: float, precision: str = "f64", walker_chunk: int = 0,
                   taylor_precision: str | None = None, vhs_layout: str = "auto",
                   taylor_hi_from: int = 0) -> CholAfqmcCtx:
    """trot/prop/chol_afqmc_ops.py:103-127 (and the host twin trot/runtime_layout.py:125-197).

    mf, sum_g mf_g L_g and 1/2 sum_g L_g L_g run on the GPU; the N x N
    exponential exp(-dt/2 h1_eff) is evaluated on the host (scipy Pade),
    as the reference's host runtime layout does (runtime_layout.py:171).
    """
    if ham_data.basis != "restricted":
        raise NotImplementedError("trot_amd propagates restricted-basis Hamiltonians")
    if not isinstance(ham_data.chol, torch.Tensor) or ham_data.chol.device.type != "cuda":
        ham_data = ham_data.to_device()
    chol = ham_data.chol
    G, N, _ = chol.shape
    dm = _total_dm(rdm1)
    mf, mfl, v0m = _lib.build_prop_tensors(chol, dm)
    h0 = float(complex(ham_data.h0 if not isinstance(ham_data.h0, torch.Tensor) else ham_data.h0.item()).real)
    mf_h = mf.cpu().numpy()
    h0_prop = complex(-h0 - 0.5 * np.sum(mf_h**2))
    h1_eff = ham_data.h1 - v0m - (1j * mfl).real  # trot/prop/chol_afqmc_ops.py:90-100
    e = scipy.linalg.expm(-0.5 * dt * h1_eff.cpu().numpy())
    ctx = CholAfqmcCtx(
        dt=float(dt),
        sqrt_dt=float(np.sqrt(dt)),
        exp_h1_half=torch.from_numpy(np.ascontiguousarray(e)).to("cuda"),
        mf_shifts=mf,
        h0_prop=h0_prop,
        chol_flat=chol.reshape(G, N * N),
        norb=N,
        precision=precision,
        taylor_precision=taylor_precision or "f64",
        walker_chunk=walker_chunk,
        taylor_hi_from=int(taylor_hi_from),
    )
    if taylor_hi_from < 0:
        raise ValueError("taylor_hi_from must be >= 0")
    if vhs_layout not in ("auto", "dense", "symblk"):
        raise ValueError(f"unknown vhs_layout: {vhs_layout}")
    # Cholesky vectors of real-orbital ERIs are symmetric; after a rotation to the MO basis only to
    # rounding.  Within 1e-12 of the largest element the sym-block path (upper triangle) applies.
    asym = big = 0.0
    for g0 in range(0, chol.shape[0], 256):  # in slices: no (G, N, N) temporary
        c = chol[g0:g0 + 256]
        asym = max(asym, float((c - c.transpose(1, 2)).abs().max()))
        big = max(big, float(c.abs().max()))
    symmetric = asym <= 1e-12 * big
    if vhs_layout == "symblk" and not symmetric:
        raise ValueError("vhs_layout='symblk' needs symmetric Cholesky vectors")
    ctx.low["symblk"] = vhs_layout != "dense" and symmetric
    if precision == "bf16x3":
        # VHS as an NT GEMM on split-bf16 planes: B = (chol columns) x G, K = G contiguous
        src = _lib.pack_symblk(chol) if ctx.low["symblk"] else ctx.chol_flat
        ctx.low["chol_b3"] = _b3.split(src, transposed=True)
        del src
    elif ctx.low["symblk"]:
        # V = sum_g xbar_g L_g is symmetric: build only its upper 8x8 blocks
        ctx.low["chol_symblk"] = _lib.pack_symblk(chol, out_f32=precision != "f64")
    elif precision != "f64":
        # the reference's mixed mode stores chol_flat as float32 (chol_afqmc_ops.py:117)
        ctx.low["chol_flat32"] = ctx.chol_flat.to(torch.float32).contiguous()
    if ctx.low["symblk"] and ctx.taylor_precision == "bf16x3":
        # whole propagation on split-BF16 MFMA: (E - I) planes for the one-body factors
        ctx.low["onebody_b3"] = _lib.onebody_b3_planes(ctx.exp_h1_half)
        # the VHS GEMM then writes V already split ([hi | lo] planes) for the propagation kernels
        ctx.low["split_v"] = precision == "bf16x3"
    return ctx


# ---------------------------------------------------------------------------
# state init (trot/prop/afqmc.py:19-76)
# ---------------------------------------------------------------------------


def _scalars(pop, e_est) -> torch.Tensor:
    return torch.tensor([float(pop), float(e_est)], dtype=torch.float64, device="cuda")


def init_prop_state(*, sys: System, ham_data, trial_ops: TrialOps, trial_data, meas_ops: MeasOps, params,
                    initial_walkers=None, initial_e_estimate=None, rdm1=None, mesh=None) -> PropState:
    """trot/prop/afqmc.py:19-76.  ``mesh`` must be None (one GPU)."""
    if mesh is not None:
        raise NotImplementedError("trot_amd runs on one GPU; mesh sharding is not supported")
    if initial_walkers is None:
        if rdm1 is None:
            rdm1 = trial_ops.get_rdm1(trial_data)
        initial_walkers = wk.init_walkers(sys=sys, rdm1=rdm1, n_walkers=params.n_walkers)
    initial_walkers = wk.as_device_walkers(initial_walkers)
    engine = meas_ops.engine
    overlaps, greens = engine.overlap_green(initial_walkers, trial_data)
    if initial_e_estimate is None:
        meas_ctx = meas_ops.build_meas_ctx(ham_data, trial_data)
        w0 = wk.take_walkers(initial_walkers, [0])
        _, g0 = engine.overlap_green(w0, trial_data)
        from .meas import _h0

        e = engine.energy_from_green(g0, _h0(ham_data), meas_ctx)
        e_est = float(e.real.mean().item())
    else:
        e_est = float(np.real(np.asarray(initial_e_estimate.cpu() if isinstance(initial_e_estimate, torch.Tensor)
                                         else initial_e_estimate)))
    scal = _scalars(e_est, e_est)
    return PropState(
        walkers=initial_walkers,
        weights=torch.ones((params.n_walkers,), dtype=torch.float64, device="cuda"),
        overlaps=overlaps,
        rng_key=(int(params.seed), 0),
        pop_control_ene_shift=scal[0],
        e_estimate=scal[1],
        node_encounters=torch.zeros((1,), dtype=torch.int64, device="cuda"),
        green=greens,
    )


# ---------------------------------------------------------------------------
# the step (trot/prop/afqmc.py:79-144)
# ---------------------------------------------------------------------------


class KernelTimer:
    """Collects (start, end) torch.cuda events around named launches on the current stream."""

    def __init__(self):
        self.events: dict[str, list] = {}

    def wrap(self, name: str, fn, *args, **kwargs):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        out = fn(*args, **kwargs)
        e.record()
        self.events.setdefault(name, []).append((s, e))
        return out

    def mean_ms(self, name: str) -> float:
        ev = self.events.get(name, [])
        return float(np.mean([s.elapsed_time(e) for s, e in ev])) if ev else float("nan")

    def reset(self):
        self.events.clear()


def _vhs(ctx: CholAfqmcCtx, xbar: torch.Tensor) -> torch.Tensor:
    """VHS in the context's precision; complex64 when the Taylor series runs in reduced precision."""
    chol = ctx.low.get("chol_symblk")  # sym-block storage when every L_g is symmetric
    if ctx.precision == "bf16x3":
        planes = ctx.low["chol_b3"]
        dtype = torch.complex128 if ctx.taylor_precision == "f64" else torch.complex64

        def fn(x):
            if ctx.low.get("split_v"):
                return _b3.gemm_nt_split(_b3.split(x), planes)
            out = torch.empty((x.shape[0], planes.rows), dtype=dtype, device=x.device)
            return _b3.gemm_nt(_b3.split(x), planes, out)

        args = (xbar,)
    elif ctx.precision == "f64":
        fn, args = _lib.vhs_build, (xbar, ctx.chol_flat if chol is None else chol)
    else:
        chol = ctx.low["chol_flat32"] if chol is None else chol
        fn, args = _lib.vhs_build_lp, (xbar, chol, ctx.precision, ctx.taylor_precision == "f64")
    timer = ctx.low.get("timer")
    if timer is not None:
        return timer.wrap("vhs", fn, *args)
    return fn(*args)


def _one_body(ctx: CholAfqmcCtx, w: torch.Tensor) -> torch.Tensor:
    """exp(-dt/2 h1_eff) w: FP64 MFMA, or FP32 MFMA when the Taylor terms run in reduced precision."""
    if ctx.taylor_precision == "f64":
        return _lib.one_body_apply(ctx.exp_h1_half, w)
    if "exp32" not in ctx.low:
        ctx.low["exp32"] = ctx.exp_h1_half.to(torch.float32).contiguous()
    return _lib.one_body_apply_lp(ctx.low["exp32"], w, "f32")


def _taylor(ctx: CholAfqmcCtx, v: torch.Tensor, t: torch.Tensor, n_terms: int) -> torch.Tensor:
    sym = ctx.low.get("symblk", False)
    if ctx.taylor_precision == "f64":
        if v.dtype != torch.complex128:
            raise ValueError("an FP64 Taylor series needs a complex128 VHS")
        if sym:
            return _lib.taylor_apply_symblk(v, t, ctx.sqrt_dt, n_terms, inplace=True)
        return _lib.taylor_apply(v.view(t.shape[0], ctx.norb, ctx.norb), t, ctx.sqrt_dt, n_terms, inplace=True)
    if v.dtype != torch.complex64:
        raise ValueError("a reduced-precision Taylor series needs a complex64 VHS")
    if sym and ctx.taylor_precision == "bf16x3":
        return _lib.taylor_apply_b3_symblk(v, t, ctx.sqrt_dt, n_terms, inplace=True)
    if sym:
        return _lib.taylor_apply_lp_symblk(v, t, ctx.sqrt_dt, n_terms, ctx.taylor_precision, inplace=True)
    return _lib.taylor_apply_lp(v.view(t.shape[0], ctx.norb, ctx.norb), t, ctx.sqrt_dt, n_terms,
                                ctx.taylor_precision, inplace=True)


def _trotter_b3(ctx: CholAfqmcCtx, v: torch.Tensor, w: torch.Tensor, n_terms: int) -> torch.Tensor:
    """One spin set through E T_K E on split-BF16 MFMA (one-body factors as w + (E - I) w)."""
    if v.dtype == torch.int32:  # split V planes
        hi_from = ctx.taylor_hi_from if ctx.taylor_hi_from > 0 else n_terms + 1
        fn = _lib.trotter_b3v
        args = (v, w.contiguous(), ctx.low["onebody_b3"], ctx.sqrt_dt, n_terms, hi_from)
    else:
        fn = _lib.trotter_b3_symblk
        args = (v, w.contiguous(), ctx.low["onebody_b3"], ctx.sqrt_dt, n_terms)
    # a recorded graph step writes the new walkers over its static input buffer (run_b3 allows w_out == w_in)
    kw = {"out": w} if ctx.low.get("walkers_in_place") and w.is_contiguous() else {}
    timer = ctx.low.get("timer")
    if timer is not None:
        return timer.wrap("trotter", fn, *args, **kw)
    return fn(*args, **kw)


def apply_trotter(walkers, xbar: torch.Tensor, ctx: CholAfqmcCtx, n_terms: int):
    """trot/prop/chol_afqmc_ops.py:198-222: E T_K(i sqrt(dt) V(xbar)) E w, V shared by both spins.

    One VHS build, then per spin: one-body half step, Taylor series, one-body
    half step (chol_afqmc_ops.py:130-131, 152-162, utils.py:20-29), in walker
    chunks of ``ctx.walker_chunk`` (0: all walkers at once).
    """
    spins = walkers if isinstance(walkers, tuple) else (walkers,)
    W = spins[0].shape[0]
    N = ctx.norb
    chunk = ctx.walker_chunk if 0 < ctx.walker_chunk < W else W
    outs = [None] * len(spins) if chunk == W else [torch.empty_like(s) for s in spins]
    for w0 in range(0, W, chunk):
        w1 = min(W, w0 + chunk)
        v = _vhs(ctx, xbar[w0:w1])
        for i, s in enumerate(spins):
            if "onebody_b3" in ctx.low:
                r = _trotter_b3(ctx, v, s[w0:w1], n_terms)
                if chunk == W:
                    outs[i] = r
                else:
                    outs[i][w0:w1] = r
                continue
            t = _one_body(ctx, s[w0:w1])
            t = _taylor(ctx, v, t, n_terms)
            r = _one_body(ctx, t)
            if chunk == W:
                outs[i] = r
            else:
                outs[i][w0:w1] = r
        del v
    return tuple(outs) if isinstance(walkers, tuple) else outs[0]


def afqmc_step(state: PropState, *, params, ham_data, trial_data, meas_ops: MeasOps, prop_ctx: CholAfqmcCtx,
               meas_ctx, fields: torch.Tensor | None = None) -> PropState:
    """One phaseless step for all walkers.  ``fields`` (W, G) overrides the provider."""
    engine = meas_ops.engine
    nw = wk.n_walkers(state.walkers)
    G = prop_ctx.chol_flat.shape[0]
    greens = state.green
    if greens is None:
        _, greens = engine.overlap_green(state.walkers, trial_data)
    key = state.rng_key
    if fields is None:
        fields, key = prop_ctx.fields(key, nw, G, state.weights.device)
    else:
        key = (key[0], key[1] + 1)
    fb = engine.force_bias_from_green(greens, meas_ctx)  # afqmc.py:95-98
    xbar, shift_term, fb_term = _lib.field_shift(fields, fb, prop_ctx.mf_shifts, prop_ctx.sqrt_dt)  # :99-103
    walkers_new = apply_trotter(state.walkers, xbar, prop_ctx, params.n_exp_terms)  # :104-106
    # inside a recorded graph the new Green's functions overwrite the state's (read only by the force bias above)
    g_out = state.green if prop_ctx.low.get("walkers_in_place") and state.green is not None else None
    overlaps_new, greens_new = engine.overlap_green(walkers_new, trial_data, out=g_out)  # :108-110
    weights = state.weights.clone()
    scal = torch.stack([state.pop_control_ene_shift, state.e_estimate]).to(torch.float64).contiguous()
    nodes = state.node_encounters.clone()
    _lib.phaseless_update(weights, state.overlaps, overlaps_new, shift_term, fb_term, scal, nodes, prop_ctx.dt,
                          prop_ctx.h0_prop, getattr(params, "weight_floor", 1.0e-3),
                          getattr(params, "weight_cap", 100.0), getattr(params, "pop_control_damping", 0.1))
    return PropState(
        walkers=walkers_new,
        weights=weights,
        overlaps=overlaps_new,
        rng_key=key,
        pop_control_ene_shift=scal[0],
        e_estimate=scal[1],
        node_encounters=nodes,
        green=greens_new,
    )


@dataclass(frozen=True)
class PropOps:
    """trot/prop/types.py:99-103."""

    init_prop_state: Callable
    build_prop_ctx: Callable
    step: Callable


def make_prop_ops(ham_basis: str, walker_kind: str, mixed_precision: bool = False, *,
                  precision: str | None = None, taylor_precision: str | None = None,
                  walker_chunk: int = 0, vhs_layout: str = "auto", taylor_hi_from: int = 0) -> PropOps:
    """trot/prop/afqmc.py:147-180.

    ``mixed_precision=True`` is the reference's mixed mode
    (trot/prop/chol_afqmc_ops.py:247-257, afqmc.py:177): xbar rounded to
    complex64 and chol_flat to float32 for the VHS GEMM only (FP32 MFMA, FP32
    accumulation), V returned as complex128, and the one-body factors and the
    Taylor series in FP64 as for complex128 walkers (chol_afqmc_ops.py:160).
    ``precision`` may name the VHS form "f64", "f32" or "bf16x3" explicitly.
    ``taylor_precision`` (default "f64" in every form) puts the one-body and
    Taylor products on FP32 ("f32") or split-BF16 ("bf16x3") MFMA instead; the
    walkers and the running Taylor sum stay complex128.
    """
    if taylor_precision not in (None, "f64", "f32", "bf16x3"):
        raise ValueError(f"unknown taylor_precision: {taylor_precision}")
    if ham_basis not in ("restricted", "generalized"):
        raise ValueError(f"unknown ham_basis: {ham_basis}")
    walker_kind = walker_kind.lower()
    if walker_kind not in ("restricted", "unrestricted", "generalized"):
        raise ValueError(f"unknown walker_kind: {walker_kind}")
    if ham_basis != "restricted" or walker_kind == "generalized":
        raise NotImplementedError(
            f"Not implemented for ham_basis={ham_basis} and walker_kind={walker_kind}")
    prec = precision or ("f32" if mixed_precision else "f64")
    if prec not in ("f64", "f32", "bf16x3"):
        raise ValueError(f"unknown precision: {prec}")

    def step(state, *, params, ham_data, trial_data, trial_ops, meas_ops, meas_ctx, prop_ctx, fields=None):
        return afqmc_step(state, params=params, ham_data=ham_data, trial_data=trial_data, meas_ops=meas_ops,
                          prop_ctx=prop_ctx, meas_ctx=meas_ctx, fields=fields)

    def _build(ham_data, rdm1, params):
        return build_prop_ctx(ham_data, rdm1, params.dt, precision=prec, walker_chunk=walker_chunk,
                              taylor_precision=taylor_precision, vhs_layout=vhs_layout,
                              taylor_hi_from=taylor_hi_from)

    return PropOps(init_prop_state=init_prop_state, build_prop_ctx=_build, step=step)
