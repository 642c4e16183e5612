"""Many independent walker sets of uneven size advanced together: one HIP graph launch per step.

The batching form for LNO-style fragments: every fragment keeps its own Hamiltonian, trial,
propagation and measurement contexts and walker state (the reference runs them one after another,
each through trot/prop/afqmc.py:82-144).  ``FragmentBatch.capture`` records one phaseless step of every
fragment into a single HIP graph (torch.cuda.CUDAGraph on the ROCm stream) over static state buffers,
with the step's auxiliary fields read from one flat device buffer; ``step`` copies the next fields into
that buffer and replays the graph.  A step of all fragments is then one field copy and one graph launch
instead of about thirty kernel launches per fragment.  For the FP64 step (the default when every fragment
runs it) each operation of the step is moreover one grouped launch over all fragments, whose
workgroups run every fragment's own problem.  Either way the result is bit-identical to stepping the
fragments one at a time: every fragment sees the same kernels on the same data in the same order.

Captured code must not synchronise with the host, allocate outside torch's graph pool or record memset
nodes; the FP64 step path (trot_amd.prop.afqmc_step with given fields) satisfies all three.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, List

import torch

from . import _lib
from .prop import PropState


@dataclass
class Fragment:
    """One walker set with everything its step needs (the arguments of PropOps.step)."""

    sys: Any
    ham: Any
    trial: Any
    trial_ops: Any
    meas_ops: Any
    prop_ops: Any
    params: Any
    prop_ctx: Any
    meas_ctx: Any
    state: PropState
    n_fields: int = field(init=False)

    def __post_init__(self):
        self.n_fields = int(self.prop_ctx.chol_flat.shape[0])

    @property
    def n_walkers(self) -> int:
        return int(self.state.weights.shape[0])

    def step(self, fields: torch.Tensor | None = None) -> PropState:
        return self.prop_ops.step(self.state, params=self.params, ham_data=self.ham, trial_data=self.trial,
                                  trial_ops=self.trial_ops, meas_ops=self.meas_ops, meas_ctx=self.meas_ctx,
                                  prop_ctx=self.prop_ctx, fields=fields)


def _tensors(state: PropState):
    """The device tensors of a state, in a fixed order (walkers and greens may be spin tuples)."""
    out = []
    for v in (state.walkers, state.weights, state.overlaps, state.pop_control_ene_shift, state.e_estimate,
              state.node_encounters, state.green):
        if v is None:
            continue
        out.extend(x for x in (v if isinstance(v, tuple) else (v,)) if x is not None)
    return out


def _static_copy(state: PropState) -> PropState:
    """A state whose tensors are fresh contiguous copies (the graph's input buffers)."""
    def cp(v):
        if v is None:
            return None
        if isinstance(v, tuple):
            return tuple(None if x is None else x.detach().clone().contiguous() for x in v)
        return v.detach().clone().contiguous()

    green = state.green
    return PropState(walkers=cp(state.walkers), weights=cp(state.weights), overlaps=cp(state.overlaps),
                     rng_key=state.rng_key, pop_control_ene_shift=cp(state.pop_control_ene_shift),
                     e_estimate=cp(state.e_estimate), node_encounters=cp(state.node_encounters), green=cp(green))


def _cost(f: Fragment) -> float:
    """Relative cost of one step of a fragment: its VHS and Taylor flops (N^2 (G + 8 K n) per walker)."""
    N = f.prop_ctx.norb
    n = sum(f.sys.nelec) if hasattr(f.sys, "nelec") else 1
    return float(f.n_walkers) * N * N * (f.n_fields + 8 * f.params.n_exp_terms * n)


def _groupable(f: Fragment) -> bool:
    """The FP64 step with sym-block V and all walkers in one chunk: the form the grouped launches cover."""
    ctx = f.prop_ctx
    eng = f.meas_ops.engine
    return (ctx.precision == "f64" and ctx.taylor_precision == "f64" and bool(ctx.low.get("symblk"))
            and ctx.walker_chunk <= 0 and eng.precision == "f64" and getattr(eng, "kind", "") in ("rr", "ru", "uu")
            and f.sys.walker_kind in ("restricted", "unrestricted") and max(f.sys.nelec) <= _lib.MAX_NOCC)


def _groupable_mixed(f: Fragment) -> bool:
    """The mixed step the grouped propagation covers: split-BF16 force bias, VHS written as split planes on
    sym-block V and the whole propagation on split-BF16 MFMA (trotter_b3v), all walkers in one chunk."""
    ctx = f.prop_ctx
    eng = f.meas_ops.engine
    return (ctx.precision == "bf16x3" and ctx.taylor_precision == "bf16x3" and bool(ctx.low.get("split_v"))
            and ctx.walker_chunk <= 0 and eng.precision == "bf16x3"
            and getattr(eng, "kind", "") in ("rr", "ru", "uu")
            and f.sys.walker_kind in ("restricted", "unrestricted") and max(f.sys.nelec) <= _lib.MAX_NOCC)


class _GroupedPlan:
    """Static buffers of every fragment and the grouped launches of one step (FP64, sym-block V).

    One launch each over all fragments and spins: the force bias from the cached Green's functions (its
    split-K partial products, then their slice sums), the shifted fields, VHS, one-body, copy of the
    first term, the Taylor terms, one-body, C^T w, inverse and determinant, Green's functions (into the
    static state), and the overlaps with the weight update.  Every launch runs, per fragment, the
    problem of the fragment's own entry point, so the step is bit-identical to stepping the fragments
    one at a time."""

    def __init__(self, frags: List[Fragment], views, tiles: dict | None = None, fused: bool = True):
        from .trial import _coeffs

        n_terms = {int(f.params.n_exp_terms) for f in frags}
        if len(n_terms) != 1:
            raise ValueError("grouped step: fragments must share n_exp_terms")
        self.n_terms = n_terms.pop()
        self.per = []
        dev = "cuda"
        vhs = _lib.Grouped("gemm", "vhs")
        # one-body products on 64 x 32 tiles for n <= 32 (same bits as 64 x 64)
        ob1 = {op: _lib.Grouped("gemm", op) for op in ("one_body", "one_body_narrow")}
        cp = _lib.Grouped("copy")
        # Taylor terms, grouped by tile shape: the autotuned shape of the fragment's (walkers, N, n) when
        # FragmentBatch.autotune ran, else 32 x 32 for n <= 32 and 32 x 64 above (every shape: the same bits)
        tay = [{op: _lib.Grouped("gemm", op) for op in _lib.TAYLOR_GOP_OF_SHAPE.values()} for _ in range(self.n_terms)]
        ob2 = {op: _lib.Grouped("gemm", op) for op in ("one_body", "one_body_narrow")}
        # the whole propagation E . T_K . E of a walker set in one workgroup per walker when its operand fits in
        # LDS (trot_fused_prop_problem); bigger sets take the launches above
        fus = _lib.Grouped("fused")
        tiles = tiles or {}
        ovl = _lib.Grouped("gemm", "overlap")
        inv = _lib.Grouped("inv")
        grn = _lib.Grouped("gemm", "green")
        # the force bias as trot_force_bias_splitk runs it per fragment: spin a (alpha), then spin b added
        # (alpha 1); one-slice problems write fb, split ones their partials, summed in slice order
        fb_a, fb_b = _lib.Grouped("gemm", "force_bias"), _lib.Grouped("gemm", "force_bias")
        sum_a, sum_b = _lib.Grouped("sum"), _lib.Grouped("sum")
        shf, upd = _lib.Grouped("shift"), _lib.Grouped("update")
        L = _lib.lib()
        for fi, f in enumerate(frags):
            ctx, st = f.prop_ctx, f.state
            W = f.n_walkers
            G = f.n_fields
            N = ctx.norb
            chol = ctx.low["chol_symblk"]
            vcols = chol.shape[1]
            b = dict(xbar=torch.empty((W, G), dtype=torch.complex128, device=dev),
                     shift=torch.empty((W,), dtype=torch.complex128, device=dev),
                     fbt=torch.empty((W,), dtype=torch.complex128, device=dev),
                     v=torch.empty((W, vcols), dtype=torch.complex128, device=dev),
                     det=[], spins=[])
            vhs.add_gemm(b["xbar"], chol, b["v"], None, W, vcols, G)
            walkers = st.walkers if isinstance(st.walkers, tuple) else (st.walkers,)
            greens = [g for g in st.green if g is not None]
            coeffs = _coeffs(f.trial)
            for si, w in enumerate(walkers):
                n = w.shape[2]
                c = coeffs[si].contiguous()
                t = torch.empty_like(w)
                t0, t1 = torch.empty_like(w), torch.empty_like(w)
                m = torch.empty((W, n, n), dtype=torch.complex128, device=dev)
                minv = torch.empty_like(m)
                det = torch.empty((W,), dtype=torch.complex128, device=dev)
                if not (fused and fus.add_fused(b["v"], ctx.exp_h1_half, w, t, W, N, n, ctx.sqrt_dt, self.n_terms)):
                    obop = "one_body_narrow" if n <= 32 else "one_body"
                    shape = tiles.get((W, N, n, True), 3 if n <= 32 else 2)
                    tayop = _lib.TAYLOR_GOP_OF_SHAPE[shape]
                    ob1[obop].add_gemm(ctx.exp_h1_half, w, t, None, W, N, n)
                    cp.add_copy(t, t0)
                    prev, nxt = t0, t1
                    for k in range(1, self.n_terms + 1):
                        tay[k - 1][tayop].add_gemm(b["v"], prev, nxt, t, W, N, n, ctx.sqrt_dt / k)
                        prev, nxt = nxt, prev
                    ob2[obop].add_gemm(ctx.exp_h1_half, t, w, None, W, N, n)
                ovl.add_gemm(c, w, m, None, W, N, n)
                inv.add_inv(m, minv, det, W, n)
                grn.add_gemm(w, minv, greens[si], None, W, N, n)
                b["det"].append(det)
                b["spins"].append((t, t0, t1, m, minv, c))
            eng = f.meas_ops.engine
            sa, sb = eng.spins(f.meas_ctx)
            b["fb"] = torch.empty((W, G), dtype=torch.complex128, device=dev)
            b["part"] = []
            spin_fb = [(greens[0], sa, 2.0 if eng.kind == "rr" else 1.0, False)]
            if eng.kind != "rr":
                spin_fb.append((greens[1], sb, 1.0, True))
            for si, (gt, sc, alpha, acc) in enumerate(spin_fb):
                Nn = gt.shape[1] * gt.shape[2]
                S = L.trot_force_bias_slices(W, G, Nn)
                _lib._check(min(S, 0), "trot_force_bias_slices")
                if S == 1:
                    (fb_a if si == 0 else fb_b).add_force_bias(gt, sc.rot_chol_t, W, G, Nn, alpha, acc, b["fb"])
                else:
                    part = torch.empty((S, W, G), dtype=torch.complex128, device=dev)
                    b["part"].append(part)
                    fb_a.add_force_bias(gt, sc.rot_chol_t, W, G, Nn, alpha, acc, part)
                    (sum_a if si == 0 else sum_b).add_sum(part, b["fb"], S, 2 * W * G, alpha, acc)
            shf.add_shift(views[fi], b["fb"], ctx.mf_shifts, ctx.sqrt_dt, W, G, b["xbar"], b["shift"], b["fbt"])
            p = f.params
            upd.add_update(b["det"][0], b["det"][1] if len(b["det"]) > 1 else None, st.overlaps, b["shift"], b["fbt"],
                           st.weights, f._scal, st.node_encounters, W, ctx.dt, ctx.h0_prop,
                           getattr(p, "weight_floor", 1.0e-3), getattr(p, "weight_cap", 100.0),
                           getattr(p, "pop_control_damping", 0.1))
            self.per.append(b)
        terms = [g for groups in tay for g in groups.values() if g.descs]
        first = [g for g in ob1.values() if g.descs]
        last = [g for g in ob2.values() if g.descs]
        copies = [cp] if cp.descs else []
        fusedl = [fus] if fus.descs else []
        self.n_fused = len(fus.descs)
        self.ops = [g.finish() for g in (vhs, *fusedl, *first, *copies, *terms, *last, ovl, inv, grn)]
        # spin a's sums write fb before spin b's one-slice GEMMs and sums add to it
        self.pre = [g.finish() for g in (fb_a, sum_a, fb_b, sum_b, shf) if g.descs]
        self.post = [upd.finish()]

    def launch_pre(self):
        for g in self.pre:
            g.launch()

    def launch(self):
        for g in self.ops:
            g.launch()

    def launch_post(self):
        for g in self.post:
            g.launch()


class _GroupedPlanMixed:
    """The grouped step for fragments in the split-BF16 mixed form (``_groupable_mixed``).

    One launch each over all fragments and spins: the split of the Green's functions into bf16 planes, the
    split-BF16 force-bias GEMM (its K slices) and the slice sums of spin a then spin b, the shifted fields,
    the split of the shifted fields, the VHS GEMM writing V as split planes, every launch of the split-BF16
    propagation (trot_trotter_b3v_stages / trot_prop_b3_grouped: the split of the walkers, the two one-body
    products and each Taylor term, walkers updated in place), C^T w, inverse and determinant, Green's
    functions, and the weight update (the FP64 parts exactly as _GroupedPlan runs them).  Each walker set runs
    the kernels and operands of its own step, so the results are bit-identical to stepping the fragments one
    at a time."""

    def __init__(self, frags: List[Fragment], views):
        from . import _b3
        from .trial import _coeffs

        n_terms = {int(f.params.n_exp_terms) for f in frags}
        if len(n_terms) != 1:
            raise ValueError("grouped step: fragments must share n_exp_terms")
        K = n_terms.pop()
        dev = "cuda"
        L = _lib.lib()
        self.per = []
        split_g, gemm_fb = _lib.Grouped("b3split"), _lib.Grouped("b3gemm", 1)
        red_a, red_b = _lib.Grouped("b3reduce"), _lib.Grouped("b3reduce")
        split_x, gemm_v = _lib.Grouped("b3split"), _lib.Grouped("b3gemm", 2)
        shf, upd = _lib.Grouped("shift"), _lib.Grouped("update")
        ovl, inv, grn = _lib.Grouped("gemm", "overlap"), _lib.Grouped("inv"), _lib.Grouped("gemm", "green")
        stage_seq, stage_groups = None, []
        for fi, f in enumerate(frags):
            ctx, st = f.prop_ctx, f.state
            W, G, N = f.n_walkers, f.n_fields, ctx.norb
            planes = ctx.low["chol_b3"]
            b = dict(xbar=torch.empty((W, G), dtype=torch.complex128, device=dev),
                     shift=torch.empty((W,), dtype=torch.complex128, device=dev),
                     fbt=torch.empty((W,), dtype=torch.complex128, device=dev),
                     fb=torch.empty((W, G), dtype=torch.complex128, device=dev),
                     v=torch.empty((2, W, planes.rows), dtype=torch.int32, device=dev), det=[], spins=[], ws=[])
            walkers = st.walkers if isinstance(st.walkers, tuple) else (st.walkers,)
            greens = [g for g in st.green if g is not None]
            # the force bias as the split-BF16 engine runs it (trot_amd.meas, force_bias_from_green): per spin,
            # the Green's function split into planes and an NT GEMM with 8 K slices; spin a's slice sums write
            # alpha * fb, spin b's add to it
            eng = f.meas_ops.engine
            sa, sb = eng.spins(f.meas_ctx)
            spin_fb = [(greens[0], sa, 2.0 if eng.kind == "rr" else 1.0, 0.0)]
            if eng.kind != "rr":
                spin_fb.append((greens[1], sb, 1.0, 1.0))
            for si, (gt, s, al, be) in enumerate(spin_fb):
                Nn = gt.shape[1] * gt.shape[2]
                if "rot_chol_t_b3" not in s.low:
                    s.low["rot_chol_t_b3"] = _b3.split(s.rot_chol_t.reshape(G, -1))
                rcb = s.low["rot_chol_t_b3"]
                gp = torch.empty(L.trot_b3_planes_bytes(W, Nn, 1), dtype=torch.uint8, device=dev)
                split_g.add_b3_split(gt.view(W, Nn), gp)
                ws = torch.empty(max(int(L.trot_gemm_b3_workspace(W, G, 1, 8)), 1), dtype=torch.uint8, device=dev)
                desc = gemm_fb.add_b3_gemm(gp, rcb.data, W, G, Nn, b["fb"], G, al, be, 8, ws)
                if not (red_a if si == 0 else red_b).add_b3_reduce(desc):
                    raise ValueError("grouped mixed step: a force bias without split K would race between spins")
                b["ws"] += [gp, ws]
            shf.add_shift(views[fi], b["fb"], ctx.mf_shifts, ctx.sqrt_dt, W, G, b["xbar"], b["shift"], b["fbt"])
            xp = torch.empty(L.trot_b3_planes_bytes(W, G, 1), dtype=torch.uint8, device=dev)
            split_x.add_b3_split(b["xbar"], xp)
            gemm_v.add_b3_gemm(xp, planes.data, W, planes.rows, G, b["v"], planes.rows, 1.0, 0.0, 1, None)
            b["ws"].append(xp)
            coeffs = _coeffs(f.trial)
            hi_from = ctx.taylor_hi_from if ctx.taylor_hi_from > 0 else K + 1
            for si, w in enumerate(walkers):
                n = w.shape[2]
                c = coeffs[si].contiguous()
                m = torch.empty((W, n, n), dtype=torch.complex128, device=dev)
                minv = torch.empty_like(m)
                det = torch.empty((W,), dtype=torch.complex128, device=dev)
                ws = torch.empty(L.trot_taylor_b3_workspace(W, N, n), dtype=torch.uint8, device=dev)
                stages = _lib.trotter_b3v_stages(b["v"], w, ctx.low["onebody_b3"], ctx.sqrt_dt, K, hi_from, w, ws)
                seq = [k for k, _, _ in stages]
                if stage_seq is None:
                    stage_seq = seq
                    stage_groups = [_lib.Grouped("b3stage", k) for k in seq]
                elif seq != stage_seq:
                    raise ValueError("grouped mixed step: walker sets need the same propagation launches "
                                     "(one walker chunk each)")
                for g, (_, wg, desc) in zip(stage_groups, stages):
                    g.add_b3_stage(desc, wg, (b["v"], w, ws, ctx.low["onebody_b3"]))
                ovl.add_gemm(c, w, m, None, W, N, n)
                inv.add_inv(m, minv, det, W, n)
                grn.add_gemm(w, minv, greens[si], None, W, N, n)
                b["det"].append(det)
                b["spins"].append((m, minv, c))
                b["ws"].append(ws)
            p = f.params
            upd.add_update(b["det"][0], b["det"][1] if len(b["det"]) > 1 else None, st.overlaps, b["shift"], b["fbt"],
                           st.weights, f._scal, st.node_encounters, W, ctx.dt, ctx.h0_prop,
                           getattr(p, "weight_floor", 1.0e-3), getattr(p, "weight_cap", 100.0),
                           getattr(p, "pop_control_damping", 0.1))
            self.per.append(b)
        if stage_seq is not None and _lib.TROT_B3_STAGE_COPY in stage_seq:
            raise ValueError("grouped mixed step: the propagation needs its one-body factors")
        self.pre = [g.finish() for g in (split_g, gemm_fb, red_a, red_b) if g.descs] + [shf.finish()]
        self.ops = ([split_x.finish(), gemm_v.finish()] + [g.finish() for g in stage_groups]
                    + [g.finish() for g in (ovl, inv, grn)])
        self.post = [upd.finish()]
        self.n_fused = 0

    def launch_pre(self):
        for g in self.pre:
            g.launch()

    def launch(self):
        for g in self.ops:
            g.launch()

    def launch_post(self):
        for g in self.post:
            g.launch()


class FragmentBatch:
    """A list of fragments stepped together through one captured HIP graph.

    ``n_streams`` independent branches: fragments are dealt to streams by decreasing cost onto the
    least-loaded stream (longest-processing-time first), each branch runs its fragments in order, and
    the graph joins the branches at the end, so fragments on different branches may run concurrently.
    Per fragment the kernels and their order are the same as in a one-by-one run.

    ``grouped`` (default: when every fragment runs the FP64 step on sym-block V, or every fragment runs the
    split-BF16 step on sym-block V): every operation of the step runs as one launch over all fragments
    (force bias and its slice sums, shifted fields, VHS, one-body, Taylor terms, overlaps, inverses, Green's
    functions, weight update; the trot_*_grouped entry points; one launch per tile shape where fragments
    differ); the results stay bit-identical to the one-by-one run.  ``autotune`` before ``capture`` picks
    every fragment's FP64 Taylor tile shape by timing.  In split-BF16 the force-bias and VHS GEMMs, the
    operand splits, the slice sums and each stage of the split-BF16 propagation (one-body, Taylor terms,
    copies; the same stage sequence in every walker set) are one launch each over all fragments.
    ``fused`` (default): the propagation of every walker set whose N x n operand fits in LDS (N <= 256,
    n <= 48) is ONE launch over all of them (trot_fused_prop_grouped: one workgroup per walker runs the
    one-body product, all Taylor terms and the last one-body product), also bit-identical.
    """

    def __init__(self, fragments: List[Fragment], n_streams: int = 4, grouped: bool | None = None,
                 fused: bool = True):
        if not fragments:
            raise ValueError("FragmentBatch needs at least one fragment")
        if n_streams < 1:
            raise ValueError("n_streams must be >= 1")
        self.fragments = fragments
        load = [0.0] * n_streams
        self.branch = [[] for _ in range(n_streams)]
        for i in sorted(range(len(fragments)), key=lambda i: -_cost(fragments[i])):
            b = min(range(n_streams), key=lambda k: load[k])
            self.branch[b].append(i)
            load[b] += _cost(fragments[i])
        self.branch = [sorted(b) for b in self.branch if b]
        self.streams = [torch.cuda.Stream() for _ in self.branch]  # default priority
        sizes = [f.n_walkers * f.n_fields for f in fragments]
        self.offsets = [0]
        for s in sizes:
            self.offsets.append(self.offsets[-1] + s)
        self.fields = torch.zeros(self.offsets[-1], dtype=torch.float64, device="cuda")
        self.graph = None
        self.steps_done = 0
        self.mixed = all(_groupable_mixed(f) for f in fragments)
        can = self.mixed or all(_groupable(f) for f in fragments)
        if grouped and not can:
            raise ValueError("grouped step: every fragment must run the FP64 step on sym-block V, or every "
                             "fragment the split-BF16 step on split V planes")
        self.grouped = can if grouped is None else bool(grouped)
        self.plan = None
        self.tiles = {}
        self.fused = bool(fused)

    def autotune(self):
        """Pick the FP64 Taylor-term tile shape of every fragment's (walkers, N, n) per spin by timing
        (trot_autotune_taylor).  Runs eagerly, before capture; the choice changes speed, not results."""
        from . import _lib

        picked = {}
        for f in self.fragments:
            sym = bool(f.prop_ctx.low.get("symblk"))
            for n in sorted(set(f.sys.nelec)):
                key = (f.n_walkers, f.prop_ctx.norb, n, sym)
                if key not in picked:
                    picked[key] = _lib.autotune_taylor(*key)[0]
        self.tiles.update(picked)  # the grouped step's Taylor launches take these shapes
        return picked

    def field_views(self):
        return [self.fields[a:b].view(f.n_walkers, f.n_fields)
                for f, a, b in zip(self.fragments, self.offsets[:-1], self.offsets[1:])]

    def pack_fields(self, per_fragment) -> torch.Tensor:
        """One flat device tensor from per-fragment (W_f, G_f) fields, in the batch's layout."""
        return torch.cat([torch.as_tensor(x, dtype=torch.float64).reshape(-1) for x in per_fragment]).to("cuda")

    def _step_one(self, i, views):
        f = self.fragments[i]
        new = f.step(views[i])
        for dst, src in zip(_tensors(f.state), _tensors(new)):
            dst.copy_(src)
        f.state = f.state._replace(rng_key=new.rng_key)

    def _step_grouped(self, views):
        self.plan.launch_pre()  # force bias, slice sums, shifted fields
        self.plan.launch()
        self.plan.launch_post()  # overlaps, weights, population shift
        for f in self.fragments:  # the fields came from the caller: one draw of the key per step
            s, c = f.state.rng_key
            f.state = f.state._replace(rng_key=(s, c + 1))

    def _make_plan(self, views):
        """Static state with the scalars as views of one (2,) buffer the weight update writes in place."""
        for f in self.fragments:
            scal = torch.stack([f.state.pop_control_ene_shift.reshape(()), f.state.e_estimate.reshape(())]).to(
                torch.float64).contiguous()
            f.state = f.state._replace(pop_control_ene_shift=scal[0], e_estimate=scal[1])
            f._scal = scal
        if self.mixed:
            self.plan = _GroupedPlanMixed(self.fragments, views)
        else:
            self.plan = _GroupedPlan(self.fragments, views, self.tiles, self.fused)
        for f, b in zip(self.fragments, self.plan.per):
            b["scal"] = f._scal

    def _step_all(self, views):
        if self.grouped:
            return self._step_grouped(views)
        main = torch.cuda.current_stream()
        for st, idx in zip(self.streams, self.branch):
            st.wait_stream(main)
            with torch.cuda.stream(st):
                for i in idx:
                    self._step_one(i, views)
        for st in self.streams:
            main.wait_stream(st)

    def capture(self, warmup_fields: torch.Tensor):
        """Record one step of every fragment.  The warm-up step (eager, on a side stream, as graph
        capture requires) advances every fragment by one step with ``warmup_fields``."""
        for f in self.fragments:
            if f.state.green is None:
                _, g = f.meas_ops.engine.overlap_green(f.state.walkers, f.trial)
                f.state = f.state._replace(green=g)
            f.state = _static_copy(f.state)
        views = self.field_views()
        if self.grouped:
            self._make_plan(views)
        self.fields.copy_(warmup_fields)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._step_all(views)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        keys = [f.state.rng_key for f in self.fragments]
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._step_all(views)
        # capture records the step without running it: the states are those after the warm-up step
        for f, k in zip(self.fragments, keys):
            f.state = f.state._replace(rng_key=k)
        self.steps_done = 1
        return self

    def step(self, fields_flat: torch.Tensor):
        """One phaseless step of every fragment: copy the fields in, replay the graph."""
        if self.graph is None:
            raise RuntimeError("FragmentBatch.step before capture")
        self.fields.copy_(fields_flat)
        self.graph.replay()
        for f in self.fragments:
            s, c = f.state.rng_key
            f.state = f.state._replace(rng_key=(s, c + 1))
        self.steps_done += 1
