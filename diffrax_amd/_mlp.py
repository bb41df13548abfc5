"""Neural-ODE path: ODETerm(MLPField(...)) with Tsit5 + PIDController on the MFMA tile
kernel (erk_mlp.hip), and its discrete adjoint (mlp_adj.hip)."""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib



def _scalar_tol(v):
    if np.ndim(v) > 0:
        raise NotImplementedError("the neural-ODE kernel takes scalar rtol / atol")
    return float(v)

class MlpProblem(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("d", ctypes.c_int32), ("hidden", ctypes.c_int32), ("n_ts", ctypes.c_int32),
                ("save_t0", ctypes.c_int32), ("save_t1", ctypes.c_int32),
                ("max_steps", ctypes.c_int32), ("ckpt_every", ctypes.c_int32),
                ("force_dtmin", ctypes.c_int32), ("batch", ctypes.c_int64),
                ("t0", ctypes.c_double), ("t1", ctypes.c_double), ("dt0", ctypes.c_double),
                ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
                ("pcoeff", ctypes.c_double), ("icoeff", ctypes.c_double),
                ("dcoeff", ctypes.c_double), ("dtmin", ctypes.c_double),
                ("dtmax", ctypes.c_double), ("factormin", ctypes.c_double),
                ("factormax", ctypes.c_double), ("safety", ctypes.c_double),
                ("error_order", ctypes.c_double), ("waves", ctypes.c_int32),
                ("save_steps", ctypes.c_int32), ("ts_stride", ctypes.c_int64),
                ("n_events", ctypes.c_int32), ("event_root_find", ctypes.c_int32),
                ("n_step_ts", ctypes.c_int32), ("n_jump_ts", ctypes.c_int32),
                ("resume", ctypes.c_int32)]


class MlpAdjIO(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("grid", "ckpt", "stats", "ts", "ys_bar", "ybar", "zs", "rows", "partial", "dw")]


class MlpBacksolveIO(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("ys", "ts_out", "ys_bar", "state", "rows", "partial", "dw", "ybar", "result")]


class MlpIO(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("y0", "ts", "ys", "ts_out", "stats", "result", "grid", "ckpt", "dense_ts",
                 "dense_ys", "dense_k", "event_coeffs", "event_flags", "event_mask", "step_ts",
                 "jump_ts", "state_in", "state_out")]


_P = ctypes.POINTER
_SIGS = {
    "dfx_mlp_pack_bytes": ([ctypes.c_int32, _P(ctypes.c_size_t)], ctypes.c_int),
    "dfx_mlp_n_out": ([_P(MlpProblem)], ctypes.c_int),
    "dfx_mlp_dense_eval": ([_P(MlpProblem), _P(_lib.DenseIO), ctypes.c_void_p], ctypes.c_int),
    "dfx_mlp_pack": ([ctypes.c_int32] + [ctypes.c_void_p] * 5, ctypes.c_int),
    "dfx_mlp_n_ckpt": ([_P(MlpProblem)], ctypes.c_int),
    "dfx_mlp_eval": ([ctypes.c_int32] + [ctypes.c_void_p] * 6 + [ctypes.c_int64, ctypes.c_void_p],
                     ctypes.c_int),
    "dfx_mlp_solve": ([_P(MlpProblem), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                       ctypes.c_void_p, _P(MlpIO), ctypes.c_void_p], ctypes.c_int),
}
_SIGS.update({
    "dfx_mlp_adj_workspace": ([_P(MlpProblem), ctypes.c_int32, _P(ctypes.c_size_t),
                               _P(ctypes.c_size_t), _P(ctypes.c_size_t)], ctypes.c_int),
    "dfx_mlp_adj_segment": ([_P(MlpProblem), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p, _P(MlpAdjIO), ctypes.c_int32, ctypes.c_int32,
                             ctypes.c_void_p], ctypes.c_int),
    "dfx_mlp_backsolve_workspace": ([_P(MlpProblem), ctypes.c_int32, _P(ctypes.c_size_t),
                                     _P(ctypes.c_size_t), _P(ctypes.c_size_t)], ctypes.c_int),
    "dfx_mlp_backsolve": ([_P(MlpProblem), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_void_p, _P(MlpBacksolveIO), ctypes.c_int32, ctypes.c_int32,
                           ctypes.c_void_p], ctypes.c_int),
    "dfx_mlp_backsolve_status": ([_P(MlpProblem), ctypes.c_int32, ctypes.c_void_p,
                                  _P(ctypes.c_void_p)], ctypes.c_int),
})
_lib._SIGS.update(_SIGS)

PRECISION = {"f32": _lib.F32, "bf16": _lib.BF16_F32}


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def pack_weights(W1, W2, W3, matmul, device):
    """MFMA fragment packing of the three weight matrices (device kernel)."""
    lib = _lib.load()
    prec = PRECISION[matmul]
    nbytes = ctypes.c_size_t(0)
    _lib.check(lib.dfx_mlp_pack_bytes(prec, ctypes.byref(nbytes)), "dfx_mlp_pack_bytes")
    packed = torch.empty(int(nbytes.value), dtype=torch.uint8, device=device)
    ws = [w.detach().to(device=device, dtype=torch.float32).contiguous() for w in (W1, W2, W3)]
    _lib.check(lib.dfx_mlp_pack(prec, *(ctypes.c_void_p(w.data_ptr()) for w in ws),
                                ctypes.c_void_p(packed.data_ptr()), _stream()), "dfx_mlp_pack")
    return packed, ws


class MlpPlan:
    """Prepared batched neural-ODE solve (launch() enqueues without host sync)."""

    def __init__(self, vfo, t0, t1, dt0, y0, ts, ctrl, sub, max_steps, ckpt_every, device,
                 dense=False, event=None, clip=None, state=None):
        lib = _lib.load()
        self.vfo = vfo
        B = y0.shape[0]
        self.y0 = y0.to(device=device, dtype=torch.float32).contiguous()
        W1, b1, W2, b2, W3, b3 = vfo.params
        self.packed, self.wf = pack_weights(W1, W2, W3, vfo.matmul, device)
        self.b = [b.detach().to(device=device, dtype=torch.float32).contiguous() for b in (b1, b2, b3)]
        n_ts = 0 if ts is None else int(ts.shape[-1])
        self.ts = (ts.to(device=device, dtype=torch.float64).contiguous() if ts is not None else None)
        ts_stride = n_ts if ts is not None and ts.ndim == 2 else 0
        nan = float("nan")
        self.prob = MlpProblem(
            solver=_lib.TSIT5, precision=PRECISION[vfo.matmul], d=64, hidden=256, n_ts=n_ts,
            save_t0=int(sub.t0), save_t1=int(sub.t1), max_steps=int(max_steps),
            ckpt_every=int(ckpt_every), force_dtmin=int(ctrl.force_dtmin), batch=B,
            t0=float(t0), t1=float(t1), dt0=nan if dt0 is None else float(dt0),
            rtol=_scalar_tol(ctrl.rtol),
            atol=_scalar_tol(ctrl.atol),
            pcoeff=float(ctrl.pcoeff), icoeff=float(ctrl.icoeff), dcoeff=float(ctrl.dcoeff),
            dtmin=nan if ctrl.dtmin is None else float(ctrl.dtmin),
            dtmax=nan if ctrl.dtmax is None else float(ctrl.dtmax),
            factormin=ctrl.factormin, factormax=ctrl.factormax, safety=ctrl.safety,
            error_order=nan if ctrl.error_order is None else float(ctrl.error_order),
            save_steps=int(sub.steps), ts_stride=ts_stride)
        self.n_out = int(lib.dfx_mlp_n_out(ctypes.byref(self.prob)))
        self.ys = torch.empty((B, self.n_out, 64), dtype=torch.float32, device=device)
        self.ts_out = torch.empty((B, self.n_out), dtype=torch.float64, device=device)
        self.stats = torch.empty((B, 3), dtype=torch.int32, device=device)
        self.result = torch.empty((B,), dtype=torch.int32, device=device)
        if ckpt_every > 0:
            n_ck = lib.dfx_mlp_n_ckpt(ctypes.byref(self.prob))
            self.grid = torch.empty((B, int(max_steps), 2), dtype=torch.float64, device=device)
            self.ckpt = torch.empty((B, n_ck, 64), dtype=torch.float32, device=device)
        else:
            self.grid = self.ckpt = None
        io = MlpIO()
        io.y0 = self.y0.data_ptr()
        io.ts = self.ts.data_ptr() if self.ts is not None else None
        io.ys, io.ts_out = self.ys.data_ptr(), self.ts_out.data_ptr()
        io.stats, io.result = self.stats.data_ptr(), self.result.data_ptr()
        io.grid = self.grid.data_ptr() if self.grid is not None else None
        io.ckpt = self.ckpt.data_ptr() if self.ckpt is not None else None
        self.event = event
        self.event_mask = None
        if event is not None:
            self.prob.n_events = len(event.conds)
            self.prob.event_root_find = int(event.root_finder is not None)
            self.event_t = dict(
                coeffs=torch.tensor([c.coeffs(64) for c in event.conds], dtype=torch.float64,
                                    device=device),
                flags=torch.tensor(event.flags(), dtype=torch.int32, device=device))
            self.event_mask = torch.zeros((B, len(event.conds)), dtype=torch.int32, device=device)
            io.event_coeffs = self.event_t["coeffs"].data_ptr()
            io.event_flags = self.event_t["flags"].data_ptr()
            io.event_mask = self.event_mask.data_ptr()
        self.clip_t = {}
        for name in ("step_ts", "jump_ts"):   # ClipStepSizeController's times, sorted
            arr = None if clip is None else getattr(clip, name)
            if arr is not None and arr.size > 0:
                self.clip_t[name] = torch.tensor(arr, dtype=torch.float64, device=device)
                setattr(self.prob, "n_" + name, int(arr.size))
                setattr(io, name, self.clip_t[name].data_ptr())
        self.state_out = None
        if state is not None:   # (resume mask, [B, 73] records or None, export wanted)
            mask, rec, want = state
            if mask:
                self.prob.resume = int(mask)
                self.state_in = rec.to(device=device, dtype=torch.float64).contiguous()
                io.state_in = self.state_in.data_ptr()
            if want:
                self.state_out = torch.empty((B, 64 + 9), dtype=torch.float64, device=device)
                io.state_out = self.state_out.data_ptr()
        self.dense = None
        if dense:
            M = int(max_steps)
            self.dense = dict(
                dense_ts=torch.empty((B, M + 1), dtype=torch.float64, device=device),
                dense_ys=torch.empty((B, M + 1, 64), dtype=torch.float32, device=device),
                dense_k=torch.empty((B, M, 7, 64), dtype=torch.float32, device=device))
            for name, v in self.dense.items():
                setattr(io, name, v.data_ptr())
        self.io = io

    def launch(self):
        lib = _lib.load()
        rc = lib.dfx_mlp_solve(ctypes.byref(self.prob), ctypes.c_void_p(self.packed.data_ptr()),
                               *(ctypes.c_void_p(b.data_ptr()) for b in self.b),
                               ctypes.byref(self.io), _stream())
        _lib.check(rc, "dfx_mlp_solve")
        self._nmax_event = None   # a copy noted for an earlier launch is stale

    def note_max_accepted(self):
        """Queue a copy of the largest accepted-step count to pinned host memory behind the
        solve; the backward pass sizes its segments from it (max_accepted) without
        draining the stream."""
        self._nmax_host = torch.empty((), dtype=torch.int32, pin_memory=True)
        self._nmax_host.copy_(self.stats[:, 1].amax(), non_blocking=True)
        self._nmax_event = torch.cuda.Event()
        self._nmax_event.record()

    def max_accepted(self):
        if getattr(self, "_nmax_event", None) is None:
            self.note_max_accepted()
        self._nmax_event.synchronize()
        return int(self._nmax_host)


class _Probe:
    pass


def _unclip(ctrl):
    """(ClipStepSizeController or None, the PIDController inside)."""
    from ._step_size_controller import ClipStepSizeController, PIDController
    if not isinstance(ctrl, ClipStepSizeController):
        return None, ctrl
    if not isinstance(ctrl.controller, PIDController):
        raise NotImplementedError("ClipStepSizeController is provided around a PIDController only")
    if ctrl.store_rejected_steps is not None or ctrl.callback_on_reject is not None:
        raise NotImplementedError("store_rejected_steps is not provided by diffrax_amd")
    return ctrl, ctrl.controller


def prepare_mlp(terms, solver, t0, t1, dt0, y0, saveat, stepsize_controller, max_steps,
                device, ckpt_every=0, allocate=True, event=None, state=None):
    from ._solver import Tsit5
    from ._step_size_controller import PIDController
    if not isinstance(solver, Tsit5):
        raise NotImplementedError("the neural-ODE kernel implements Tsit5")
    clip, stepsize_controller = _unclip(stepsize_controller)
    if not isinstance(stepsize_controller, PIDController):
        raise NotImplementedError("the neural-ODE kernel implements PIDController")
    c = stepsize_controller
    from ._step_size_controller import rms_norm
    if c.norm is not rms_norm:
        raise NotImplementedError("the forward solve's controller uses rms_norm; "
                                  "adjoint_rms_seminorm is for BacksolveAdjoint's controller")
    sub = saveat.subs
    if sub is None:
        if not saveat.dense:
            raise ValueError("SaveAt must save something")
        from ._saveat import SubSaveAt, save_y
        sub = SubSaveAt.__new__(SubSaveAt)   # dense output only: no saved rows
        sub.t0 = sub.t1 = False
        sub.ts, sub.steps, sub.fn = None, 0, save_y
    if not hasattr(sub, "fn"):
        raise NotImplementedError("the neural-ODE kernel takes a single SubSaveAt")
    if sub.fn is not save_y_fn():
        raise NotImplementedError("SaveAt(fn=...) is provided on the small-system kernel")
    if ckpt_every > 0 and (sub.steps or saveat.dense):
        raise NotImplementedError("gradients through SaveAt(steps=...) or SaveAt(dense=True) "
                                  "are not provided on the neural-ODE kernel")
    y0 = torch.as_tensor(y0)
    batched = y0.ndim == 2
    y0b = y0 if batched else y0.reshape(1, -1)
    ts = None
    check = None
    if sub.ts is not None:
        # device-resident ts are checked from a pinned copy once the solve is queued
        from ._integrate import _HostCopies
        ts = torch.as_tensor(sub.ts, dtype=torch.float64)
        if ts.ndim == 2 and ts.shape[0] == y0b.shape[0]:
            pass   # per-trajectory save times [B, n_ts]
        elif ts.ndim <= 1:
            ts = ts.reshape(-1)
        else:
            raise ValueError(f"SaveAt(ts) must be [n] or [{y0b.shape[0]}, n]")
        host = _HostCopies()
        host.add("ts", ts)
        host.seal()
        t0f, t1f = float(t0), float(t1)

        def check():
            dirn = 1.0 if t0f < t1f else -1.0
            tn = host.get("ts") * dirn
            if tn.shape[-1] > 1 and (tn[..., 1:] < tn[..., :-1]).any():
                raise RuntimeError("saveat.ts must be increasing or decreasing.")
            if ((tn > t1f * dirn) | (tn < t0f * dirn)).any():
                raise RuntimeError("saveat.ts must lie between t0 and t1.")
        if not host.deferred:
            check()
            check = None
    if not allocate:
        probe = _Probe()
        probe.ts_host = ts
        probe.deferred_checks = check
        probe.sub = sub
        return probe
    plan = MlpPlan(terms.vector_field, t0, t1, dt0, y0b, ts, c, sub, max_steps, ckpt_every, device,
                   dense=saveat.dense, event=event, clip=clip, state=state)
    plan.batched = batched
    plan.deferred_checks = check
    return plan


_ = (math, np)


def _state_outputs(plan, saveat):
    """Solution.solver_state / controller_state / made_jump from the final state records
    (_integrate.py:1489-1500; the layout of DFX_MLP_STATE_WORDS), batch leading."""
    rec = getattr(plan, "state_out", None)
    if rec is None or saveat is None:
        return None, None, None
    d = 64
    pick = (lambda v: v) if plan.batched else (lambda v: v[0])
    solver_state = controller_state = made_jump = None
    if saveat.solver_state:
        solver_state = (pick(rec[:, d] != 0), pick(rec[:, :d].to(torch.float32)))
    if saveat.controller_state:
        controller_state = (pick(rec[:, d + 2]), pick(rec[:, d + 3]), pick(rec[:, d + 4] != 0))
    if saveat.made_jump:
        made_jump = pick(rec[:, d + 1] != 0)
    return solver_state, controller_state, made_jump


def _solution(plan, t0, t1, max_steps, throw, ys, saveat=None):
    from ._solution import RESULTS, Solution
    if getattr(plan, "deferred_checks", None) is not None:
        plan.deferred_checks()
    st = plan.stats
    stats = {"num_steps": st[:, 0], "num_accepted_steps": st[:, 1],
             "num_rejected_steps": st[:, 2], "max_steps": max_steps}
    ts, result = plan.ts_out, plan.result
    if not plan.batched:
        ys, ts, result = ys[0], ts[0], result[0]
        stats = {k: (v[0] if isinstance(v, torch.Tensor) else v) for k, v in stats.items()}
    if throw:
        bad = (plan.result != 0) & (plan.result != int(RESULTS.event_occurred))   # is_okay
        if bool(bad.any()):
            raise RuntimeError(RESULTS[int(plan.result[bad][0].item())])
    event_mask = None
    if getattr(plan, "event_mask", None) is not None:
        m = plan.event_mask.bool()
        if not plan.batched:
            m = m[0]
        event_mask = ([m[..., i] for i in range(m.shape[-1])] if plan.event.is_list
                      else m[..., 0])
    interp = None
    if getattr(plan, "dense", None) is not None:
        interp = MlpDenseInterpolation(plan)
        p = plan.prob
        if p.n_ts == 0 and not (p.save_t0 or p.save_t1 or p.save_steps):
            ts = ys = None   # SaveAt(dense=True) alone saves no rows
    solver_state, controller_state, made_jump = _state_outputs(plan, saveat)
    return Solution(t0=t0, t1=t1, ts=ts, ys=ys, stats=stats, result=result,
                    interpolation=interp, event_mask=event_mask, solver_state=solver_state,
                    controller_state=controller_state, made_jump=made_jump)


def save_y_fn():
    from ._saveat import save_y
    return save_y


class MlpDenseInterpolation:
    """Dense output of a neural-ODE SaveAt(dense=True) solve (DenseInterpolation,
    _global_interpolation.py:315-400), evaluated on the GPU by dfx_mlp_dense_eval.
    ``evaluate(t0, t1=None, left=True)`` and ``derivative(t, left=True)`` take a scalar
    time, [n] times shared by the batch or [B, n] per-trajectory times."""

    def __init__(self, plan):
        self.plan = plan

    def evaluate(self, t0, t1=None, left: bool = True):
        if t1 is not None:
            return self.evaluate(t1, left=left) - self.evaluate(t0, left=left)
        return self._query(t0, False, left)

    def derivative(self, t, left: bool = True):
        return self._query(t, True, left)

    def _query(self, t, derivative, left):
        plan = self.plan
        dev = plan.y0.device
        B = plan.y0.shape[0]
        tq = torch.as_tensor(t, dtype=torch.float64)
        scalar = tq.ndim == 0
        if scalar:
            tq = tq.reshape(1)
        if tq.ndim == 1:
            n_q, stride = tq.shape[0], 0
        elif tq.ndim == 2 and tq.shape[0] == B:
            n_q, stride = tq.shape[1], tq.shape[1]
        else:
            raise ValueError(f"evaluation times must be a scalar, [n] or [{B}, n]")
        tq = tq.to(dev).contiguous()
        out = torch.empty((B, n_q, 64), dtype=torch.float32, device=dev)
        D = plan.dense
        io = _lib.DenseIO(dense_ts=D["dense_ts"].data_ptr(), dense_ys=D["dense_ys"].data_ptr(),
                          dense_k=D["dense_k"].data_ptr(), stats=plan.stats.data_ptr(),
                          t1=None, tq=tq.data_ptr(), n_q=n_q, tq_stride=stride,
                          left=int(bool(left)), derivative=int(bool(derivative)),
                          out=out.data_ptr())
        _lib.check(_lib.load().dfx_mlp_dense_eval(ctypes.byref(plan.prob), ctypes.byref(io),
                                                  _stream()), "dfx_mlp_dense_eval")
        if scalar:
            out = out[:, 0]
        if not plan.batched:
            out = out[0]
        return out


class _Params:
    """The parameter set MlpPlan reads (an MLPField or the detached autograd inputs)."""

    def __init__(self, params, matmul):
        self.params = params
        self.matmul = matmul


def backward_mlp(plan, ys_bar):
    """Gradients (y0, W1, b1, W2, b2, W3, b3) of sum(ys * ys_bar) through the recorded solve."""
    lib = _lib.load()
    dev = plan.ys.device
    B = plan.ys.shape[0]
    p = plan.prob
    n_max = plan.max_accepted() if B else 0
    ys_bar = ys_bar.to(device=dev, dtype=torch.float32).contiguous()
    n_ts = p.n_ts
    ybar = (ys_bar[:, -1, :].clone() if p.save_t1 else torch.zeros((B, 64), dtype=torch.float32, device=dev))
    dw = torch.zeros(98880, dtype=torch.float32, device=dev)
    C = p.ckpt_every
    if n_max > 0:
        Cs = min(C, n_max)
        zs_b, rows_b, part_b = (ctypes.c_size_t(0) for _ in range(3))
        _lib.check(lib.dfx_mlp_adj_workspace(ctypes.byref(p), Cs, ctypes.byref(zs_b),
                                             ctypes.byref(rows_b), ctypes.byref(part_b)),
                   "dfx_mlp_adj_workspace")
        zs = torch.empty(int(zs_b.value) // 4, dtype=torch.float32, device=dev)
        rows = torch.zeros(int(rows_b.value), dtype=torch.uint8, device=dev)
        part = torch.empty(int(part_b.value) // 4, dtype=torch.float32, device=dev)
        io = MlpAdjIO(grid=plan.grid.data_ptr(), ckpt=plan.ckpt.data_ptr(),
                      stats=plan.stats.data_ptr(),
                      ts=plan.ts.data_ptr() if plan.ts is not None else None,
                      ys_bar=ys_bar.data_ptr(), ybar=ybar.data_ptr(), zs=zs.data_ptr(),
                      rows=rows.data_ptr(), partial=part.data_ptr(), dw=dw.data_ptr())
        nseg = (n_max + C - 1) // C
        for s in reversed(range(nseg)):
            _lib.check(lib.dfx_mlp_adj_segment(ctypes.byref(p), ctypes.c_void_p(plan.packed.data_ptr()),
                                               *(ctypes.c_void_p(b.data_ptr()) for b in plan.b),
                                               ctypes.byref(io), s, Cs, _stream()),
                       "dfx_mlp_adj_segment")
    y0_bar = ybar
    if p.save_t0:
        y0_bar = y0_bar + ys_bar[:, 0, :]
    if float(p.t0) == float(p.t1) and n_ts > 0:
        off = int(p.save_t0)
        y0_bar = y0_bar + ys_bar[:, off:off + n_ts, :].sum(1)
    o = 0
    shapes = [(256, 64), (256, 256), (64, 256)]
    Ws = []
    for s_ in shapes:
        n = s_[0] * s_[1]
        Ws.append(dw[o:o + n].view(*s_))
        o += n
    db1, db2, db3 = dw[o:o + 256], dw[o + 256:o + 512], dw[o + 512:o + 576]
    return y0_bar, Ws[0], db1, Ws[1], db2, Ws[2], db3


class _MlpSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y0, W1, b1, W2, b2, W3, b3, cfg):
        params = _Params([t.detach() for t in (W1, b1, W2, b2, W3, b3)], cfg["matmul"])
        need = any(ctx.needs_input_grad[:7])
        clip, ctrl = _unclip(cfg["ctrl"])
        plan = MlpPlan(params, cfg["t0"], cfg["t1"], cfg["dt0"], y0.detach(), cfg["ts"],
                       ctrl, cfg["sub"], cfg["max_steps"],
                       cfg["ckpt_every"] if need else 0, cfg["device"], clip=clip)
        set_waves(plan, cfg.get("launch"))
        plan.launch()
        plan.note_max_accepted()
        ctx.plan = plan
        ctx.dtypes = [t.dtype for t in (y0, W1, b1, W2, b2, W3, b3)]
        ctx.mark_non_differentiable(plan.stats, plan.result, plan.ts_out)
        return plan.ys, plan.stats, plan.result, plan.ts_out

    @staticmethod
    def backward(ctx, ys_bar, *_):
        grads = backward_mlp(ctx.plan, ys_bar)
        out = [g.to(dt) if need else None
               for g, dt, need in zip(grads, ctx.dtypes, ctx.needs_input_grad[:7])]
        return (*out, None)


# Lock-step iterations per backsolve launch: the row buffers hold chunk * 7 stages rows of
# every trajectory (4.6 KB each in float32), and the host checks for unfinished rows after
# every launch.
BACKSOLVE_CHUNK = 8


def backsolve_problem(plan, adjoint):
    """The problem of the backward solves: the forward's, with BacksolveAdjoint's kwargs
    applied as the reference applies them (_adjoint.py:623-632).  The neural-ODE kernel
    solves them with the adjoint seminorm (_misc.py:85-116), Tsit5 and integral control."""
    from ._solver import Tsit5
    from ._step_size_controller import PIDController, adjoint_rms_seminorm
    kw = adjoint.kwargs
    ctrl = kw.get("stepsize_controller")
    if not (isinstance(ctrl, PIDController) and ctrl.norm is adjoint_rms_seminorm):
        raise NotImplementedError(
            "BacksolveAdjoint on the neural-ODE kernel solves with the adjoint seminorm: pass "
            "BacksolveAdjoint(stepsize_controller=PIDController(rtol, atol, "
            "norm=adjoint_rms_seminorm)).  Under the default norm every trajectory's "
            "98,880 weight adjoints would enter its error control, which is not provided.")
    if not isinstance(kw.get("solver", Tsit5()), Tsit5):
        raise NotImplementedError("the neural-ODE backsolve runs Tsit5")
    if float(ctrl.pcoeff) != 0.0 or float(ctrl.dcoeff) != 0.0:
        raise NotImplementedError("the neural-ODE backsolve implements the integral controller "
                                  "(pcoeff = dcoeff = 0)")
    if ("dt0" in kw and kw["dt0"] is None) or ("dt0" not in kw and math.isnan(plan.prob.dt0)):
        raise NotImplementedError("the neural-ODE backsolve needs a dt0: pass it to diffeqsolve "
                                  "or to BacksolveAdjoint")
    nan = float("nan")
    f = plan.prob
    back = MlpProblem()
    ctypes.memmove(ctypes.byref(back), ctypes.byref(f), ctypes.sizeof(f))
    back.precision = _lib.F32   # float32 MFMA for the backward products (both matmul forms)
    back.ckpt_every = 0
    back.dt0 = float(kw["dt0"]) if "dt0" in kw else -float(f.dt0)
    back.max_steps = int(kw.get("max_steps", f.max_steps))
    back.rtol, back.atol = _scalar_tol(ctrl.rtol), _scalar_tol(ctrl.atol)
    back.pcoeff, back.icoeff, back.dcoeff = float(ctrl.pcoeff), float(ctrl.icoeff), float(ctrl.dcoeff)
    back.dtmin = nan if ctrl.dtmin is None else float(ctrl.dtmin)
    back.dtmax = nan if ctrl.dtmax is None else float(ctrl.dtmax)
    back.force_dtmin = int(ctrl.force_dtmin)
    back.factormin, back.factormax, back.safety = ctrl.factormin, ctrl.factormax, ctrl.safety
    back.error_order = nan if ctrl.error_order is None else float(ctrl.error_order)
    return back


def backsolve_mlp(plan, back, ys_bar, chunk=BACKSOLVE_CHUNK):
    """Gradients (y0, W1, b1, W2, b2, W3, b3) of sum(ys * ys_bar) by the continuous adjoint
    (dfx_mlp_backsolve), and the backward solves' result codes.  The weights pack for
    float32 products (the backward runs them in float32 MFMA for either matmul form)."""
    lib = _lib.load()
    dev = plan.ys.device
    B = plan.ys.shape[0]
    ys_bar = ys_bar.to(device=dev, dtype=torch.float32).contiguous()
    ybar = torch.zeros((B, 64), dtype=torch.float32, device=dev)
    result = torch.zeros((B,), dtype=torch.int32, device=dev)
    dw = torch.zeros(98880, dtype=torch.float32, device=dev)
    if B == 0:
        return (ybar,) + _split_dw(dw), result
    W1, _, W2, _, W3, _ = plan.vfo.params
    packed, _ = pack_weights(W1, W2, W3, "f32", dev)
    st_b, rows_b, part_b = (ctypes.c_size_t(0) for _ in range(3))
    _lib.check(lib.dfx_mlp_backsolve_workspace(ctypes.byref(back), chunk, ctypes.byref(st_b),
                                               ctypes.byref(rows_b), ctypes.byref(part_b)),
               "dfx_mlp_backsolve_workspace")
    state = torch.zeros(int(st_b.value), dtype=torch.uint8, device=dev)
    rows = torch.zeros(int(rows_b.value) // 4, dtype=torch.float32, device=dev)
    part = torch.empty(int(part_b.value) // 4, dtype=torch.float32, device=dev)
    io = MlpBacksolveIO(ys=plan.ys.data_ptr(), ts_out=plan.ts_out.data_ptr(),
                        ys_bar=ys_bar.data_ptr(), state=state.data_ptr(), rows=rows.data_ptr(),
                        partial=part.data_ptr(), dw=dw.data_ptr(), ybar=ybar.data_ptr(),
                        result=result.data_ptr())
    cptr = ctypes.c_void_p(0)
    _lib.check(lib.dfx_mlp_backsolve_status(ctypes.byref(back), chunk,
                                            ctypes.c_void_p(state.data_ptr()), ctypes.byref(cptr)),
               "dfx_mlp_backsolve_status")
    off = int(cptr.value) - state.data_ptr()
    counters = state[off:off + 8].view(torch.int32)
    host = torch.empty(2, dtype=torch.int32, pin_memory=True)
    first = 1
    while True:
        _lib.check(lib.dfx_mlp_backsolve(ctypes.byref(back), ctypes.c_void_p(packed.data_ptr()),
                                         *(ctypes.c_void_p(b.data_ptr()) for b in plan.b),
                                         ctypes.byref(io), chunk, first, _stream()),
                   "dfx_mlp_backsolve")
        first = 0
        host.copy_(counters, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        if int(host[0]) == 0:
            break
    return (ybar,) + _split_dw(dw), result


def _split_dw(dw):
    """dW1 | dW2 | dW3 | db1 | db2 | db3 views of the flat accumulator."""
    o = 0
    Ws = []
    for r, c in ((256, 64), (256, 256), (64, 256)):
        Ws.append(dw[o:o + r * c].view(r, c))
        o += r * c
    return Ws[0], dw[o:o + 256], Ws[1], dw[o + 256:o + 512], Ws[2], dw[o + 512:o + 576]


class _MlpBacksolve(torch.autograd.Function):
    """ys of a neural-ODE solve with the continuous adjoint (dfx_mlp_backsolve) as its
    backward.  The backward solves restart from the returned ys, which are saved with
    save_for_backward: an in-place change of them makes backward raise."""

    @staticmethod
    def forward(ctx, y0, W1, b1, W2, b2, W3, b3, cfg):
        params = _Params([t.detach() for t in (W1, b1, W2, b2, W3, b3)], cfg["matmul"])
        plan = MlpPlan(params, cfg["t0"], cfg["t1"], cfg["dt0"], y0.detach(), cfg["ts"],
                       cfg["ctrl"], cfg["sub"], cfg["max_steps"], 0, cfg["device"])
        set_waves(plan, cfg.get("launch"))
        plan.launch()
        ctx.save_for_backward(plan.ys)
        ctx.plan = plan
        ctx.back = backsolve_problem(plan, cfg["adjoint"])
        ctx.throw = cfg["throw"]
        ctx.dtypes = [t.dtype for t in (y0, W1, b1, W2, b2, W3, b3)]
        ctx.mark_non_differentiable(plan.stats, plan.result, plan.ts_out)
        return plan.ys, plan.stats, plan.result, plan.ts_out

    @staticmethod
    def backward(ctx, ys_bar, *_):
        from ._solution import RESULTS
        ctx.saved_tensors   # the version check of the saved ys
        ys_bar = torch.zeros_like(ctx.plan.ys) if ys_bar is None else ys_bar
        grads, result = backsolve_mlp(ctx.plan, ctx.back, ys_bar)
        if ctx.throw:
            bad = result != 0
            if bool(bad.any()):
                raise RuntimeError(RESULTS[int(result[bad][0].item())])
        out = [g.to(dt) if need else None
               for g, dt, need in zip(grads, ctx.dtypes, ctx.needs_input_grad[:7])]
        return (*out, None)


def checkpoint_every(adjoint, max_steps):
    """Steps between stored states: ceil(max_steps / checkpoints), or sqrt(max_steps) when
    RecursiveCheckpointAdjoint(checkpoints=None)."""
    k = getattr(adjoint, "checkpoints", None)
    if k is None:
        return max(1, int(math.ceil(math.sqrt(max_steps))))
    return max(1, int(math.ceil(max_steps / int(k))))


def set_waves(plan, launch):
    """The tile's wave count from diffeqsolve's ``launch``: None keeps the library default
    (4 for float32, 8 for the three-piece bf16 form), 4 or 8 selects it, "auto" times both
    once per problem signature and device (_autotune.tune_mlp).  The results are the same
    bits either way."""
    if launch is None:
        return
    if launch == "auto":
        from . import _autotune
        _autotune.tune_mlp(plan)
    elif launch in (4, 8):
        plan.prob.waves = int(launch)
    else:
        raise ValueError("launch for the neural-ODE kernel is None, 'auto', 4 or 8 "
                         "(wavefronts per 32-trajectory tile)")


def diffeqsolve_mlp(terms, solver, t0, t1, dt0, y0, saveat, stepsize_controller, adjoint,
                    max_steps, throw, device, launch=None, event=None, solver_state=None,
                    controller_state=None, made_jump=None):
    vfo = terms.vector_field
    y0t = torch.as_tensor(y0)
    needs_grad = torch.is_grad_enabled() and (
        y0t.requires_grad or any(p.requires_grad for p in vfo.params))
    if event is not None and needs_grad:
        raise NotImplementedError("gradients through a neural-ODE solve with events are not "
                                  "provided")
    from ._adjoint import BacksolveAdjoint
    if (needs_grad and isinstance(adjoint, BacksolveAdjoint)
            and _unclip(stepsize_controller)[0] is not None):
        raise NotImplementedError("BacksolveAdjoint with ClipStepSizeController is not provided")
    resuming = solver_state is not None or controller_state is not None or made_jump is not None
    want_state = bool(saveat.solver_state or saveat.controller_state or saveat.made_jump)
    state = None
    if resuming or want_state:
        if needs_grad:
            raise NotImplementedError("gradients through solver / controller state export or "
                                      "resume are not provided on the neural-ODE kernel")
        clip, ctrl = _unclip(stepsize_controller)
        if resuming and clip is not None:
            raise NotImplementedError("resuming a ClipStepSizeController solve is not provided")
        from ._integrate import _state_record
        B = y0t.shape[0] if y0t.ndim == 2 else 1
        mask, rec = (0, None)
        if resuming:
            mask, rec = _state_record(solver, ctrl, B, 64, y0t.ndim == 2, solver_state,
                                      controller_state, made_jump)
        state = (mask, rec, want_state)
    if not needs_grad:
        plan = prepare_mlp(terms, solver, t0, t1, dt0, y0, saveat, stepsize_controller,
                           max_steps, device, event=event, state=state)
        set_waves(plan, launch)
        plan.launch()
        return _solution(plan, t0, t1, max_steps, throw, plan.ys, saveat)
    from ._adjoint import BacksolveAdjoint, RecursiveCheckpointAdjoint
    if not isinstance(adjoint, (RecursiveCheckpointAdjoint, BacksolveAdjoint)):
        raise NotImplementedError("gradients are provided through RecursiveCheckpointAdjoint "
                                  "and BacksolveAdjoint")
    # validation and shapes via a throwaway-free preparation of the configuration
    probe = prepare_mlp(terms, solver, t0, t1, dt0, y0t.detach(), saveat, stepsize_controller,
                        max_steps, device, ckpt_every=0, allocate=False)
    if probe.sub.steps or saveat.dense:
        raise NotImplementedError("gradients through SaveAt(steps=...) or SaveAt(dense=True) "
                                  "are not provided on the neural-ODE kernel")
    cfg = dict(matmul=vfo.matmul, t0=t0, t1=t1, dt0=dt0, ts=probe.ts_host, ctrl=stepsize_controller,
               sub=saveat.subs, max_steps=max_steps, device=device, launch=launch)
    batched = y0t.ndim == 2
    y0b = y0t if batched else y0t.reshape(1, -1)
    if isinstance(adjoint, BacksolveAdjoint):
        cfg.update(adjoint=adjoint, throw=bool(adjoint.kwargs.get("throw", throw)))
        ys, stats, result, ts_out = _MlpBacksolve.apply(y0b.to(device), *vfo.params, cfg)
    else:
        cfg.update(ckpt_every=checkpoint_every(adjoint, max_steps))
        ys, stats, result, ts_out = _MlpSolve.apply(y0b.to(device), *vfo.params, cfg)

    class _P:
        pass
    pl = _P()
    pl.stats, pl.result, pl.ts_out, pl.batched = stats, result, ts_out, batched
    pl.deferred_checks = probe.deferred_checks
    return _solution(pl, t0, t1, max_steps, throw, ys)


def mlp_eval(vfo, y):
    """Evaluate the registered MLP vector field on the GPU for states y [R, 64]."""
    lib = _lib.load()
    dev = y.device
    W1, b1, W2, b2, W3, b3 = vfo.params
    packed, _ = pack_weights(W1, W2, W3, vfo.matmul, dev)
    bs = [b.detach().to(device=dev, dtype=torch.float32).contiguous() for b in (b1, b2, b3)]
    y = y.to(torch.float32).contiguous()
    f = torch.empty_like(y)
    _lib.check(lib.dfx_mlp_eval(PRECISION[vfo.matmul], ctypes.c_void_p(packed.data_ptr()),
                                *(ctypes.c_void_p(b.data_ptr()) for b in bs),
                                ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(f.data_ptr()),
                                y.shape[0], _stream()), "dfx_mlp_eval")
    return f
