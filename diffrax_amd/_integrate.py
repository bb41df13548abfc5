"""diffeqsolve (mirrors diffrax/_integrate.py:888-1543) on MI355X kernels.

The Python layer does what the reference does at trace time -- argument checking
(_integrate.py:1036-1235), dtype promotion, output allocation (1272-1305) -- and then
hands the whole integration loop to one persistent HIP kernel through the C ABI.  The
batch axis that jax.vmap would add is explicit: y0 of shape [B, d] solves B instances,
with t1, SaveAt(ts) and vector-field args either shared or given per instance."""
from __future__ import annotations

import ctypes
import math
from typing import Any, Optional

import numpy as np
import torch

from . import _lib
from ._adjoint import AbstractAdjoint, BacksolveAdjoint, RecursiveCheckpointAdjoint
from ._saveat import SaveAt, SubSaveAt, save_y
from ._solution import RESULTS, Solution
from ._solver import AbstractERK, AbstractSolver
from ._step_size_controller import (ClipStepSizeController, ConstantStepSize, PIDController,
                                    rms_norm)
from ._term import ControlTerm, IncrementPath, MultiTerm, ODETerm
from ._vector_fields import DiagonalNoise, MLPField, OUDrift, RegisteredVectorField

_SMALL_VFS = (_lib.VF_LV, _lib.VF_VDP, _lib.VF_LINEAR)


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("diffrax_amd runs on a ROCm GPU (MI355X); no GPU is visible and "
                           "there is no CPU fallback.")
    return torch.device("cuda", torch.cuda.current_device())


def _nan_if_none(x):
    return float("nan") if x is None else float(x)


def _scalar_time(x, name):
    if isinstance(x, torch.Tensor):
        if x.numel() != 1:
            return None
        return float(x.reshape(()).item())
    if isinstance(x, np.ndarray):
        if x.size != 1:
            return None
        return float(x.reshape(()))
    return float(x)


class _HostCopies:
    """Host values of device-resident inputs that the host checks (t1, SaveAt ts).

    A CPU tensor is read in place.  A device tensor is copied to pinned host memory behind
    the work already queued on the current stream, so the solve can be queued before the
    checks run: ``get`` then waits for those copies only, never for the solve."""

    def __init__(self):
        self._vals = {}
        self._event = None
        self._pending = False

    def add(self, name, t):
        if t.device.type == "cpu":
            self._vals[name] = t.numpy()
            return
        h = torch.empty(tuple(t.shape), dtype=t.dtype, pin_memory=True)
        h.copy_(t, non_blocking=True)
        self._vals[name] = h
        self._pending = True

    @property
    def deferred(self):
        return self._pending

    def seal(self):
        if self._pending:
            self._event = torch.cuda.Event()
            self._event.record()

    def get(self, name):
        if self._event is not None:
            self._event.synchronize()
        v = self._vals[name]
        return v.numpy() if isinstance(v, torch.Tensor) else v


class _Workspace:
    """Per-device scratch buffers owned by the caller side (torch allocator)."""

    def __init__(self):
        self._buf = {}

    def get(self, nbytes, device):
        key = (device.index, "ws")
        buf = self._buf.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
            self._buf[key] = buf
        return buf


_WS = _Workspace()


class ErkPlan:
    """A fully prepared batched ERK solve: inputs, outputs and the C structs.

    ``launch()`` enqueues the solve on the current stream (no host sync), so a plan can be
    replayed or captured into a CUDA/HIP graph by the caller."""

    def __init__(self, problem, tensors, n_out, batched, launch_cfg=None):
        self.problem = problem
        self.t = tensors
        self.n_out = n_out
        self.batched = batched
        self.launch_cfg = launch_cfg
        self.deferred_checks = None
        io = _lib.ErkIO()
        self.io = io
        self._bind()
        nbytes = ctypes.c_size_t(0)
        lib = _lib.load()
        _lib.check(lib.dfx_erk_workspace_bytes(ctypes.byref(problem), ctypes.byref(nbytes)),
                   "dfx_erk_workspace_bytes")
        self.ws_bytes = int(nbytes.value)
        self.ws = torch.empty(max(self.ws_bytes, 256), dtype=torch.uint8,
                              device=tensors["y0"].device)

    def _bind(self):
        for name in ("y0", "t1", "ts", "params", "ys", "ts_out", "stats", "result", "grid",
                     "dense_ts", "dense_ys", "dense_k", "step_ts", "jump_ts",
                     "event_coeffs", "event_flags", "event_mask", "grid_y", "rtol", "atol",
                     "ckpt", "state_in", "state_out"):
            v = self.t.get(name)
            setattr(self.io, name, v.data_ptr() if v is not None else None)

    def record_grid(self):
        """Record every accepted step's (t_start, t_end) and start state ([B, max_steps, 2]
        and [B, max_steps, d]; a diagnostic output, not needed by the adjoint)."""
        B, d = self.t["y0"].shape
        M = int(self.problem.max_steps)
        dev = self.t["y0"].device
        if self.t.get("grid") is None:
            self.t["grid"] = torch.empty((B, M, 2), dtype=torch.float64, device=dev)
        self.t["grid_y"] = torch.empty((B, M, d), dtype=self.t["y0"].dtype, device=dev)
        self._bind()

    def record_checkpoints(self, checkpoints=None):
        """Store the lane state every K accepted steps, the restart points of the discrete
        adjoint: K = ceil(max_steps / checkpoints), checkpoints = ceil(sqrt(max_steps)) when
        None (RecursiveCheckpointAdjoint.checkpoints, _adjoint.py:264-329).  Memory is
        [B, ceil(max_steps / K), d + 8] float64 here and [B, K] records of (2 + d) values in
        the adjoint's workspace."""
        M = max(int(self.problem.max_steps), 1)
        c = int(math.ceil(math.sqrt(M))) if checkpoints is None else int(checkpoints)
        if c < 1:
            raise ValueError("RecursiveCheckpointAdjoint(checkpoints=...) must be at least 1")
        self.problem.ckpt_every = -(-M // min(c, M))
        n_ck, words = ctypes.c_int32(0), ctypes.c_int32(0)
        lib = _lib.load()
        _lib.check(lib.dfx_erk_ckpt_layout(ctypes.byref(self.problem), ctypes.byref(n_ck),
                                           ctypes.byref(words)), "dfx_erk_ckpt_layout")
        B = self.t["y0"].shape[0]
        self.t["ckpt"] = torch.empty((B, n_ck.value, words.value), dtype=torch.float64,
                                     device=self.t["y0"].device)
        self._bind()

    def adjoint(self, ys_bar, y0_bar, params_bar, workspace=None, stream=None):
        """dfx_erk_adjoint over this plan's checkpoints (after launch()): cotangents of
        the saved rows in, per-instance y0 and parameter cotangents out.  ``workspace``
        may be a uint8 tensor of at least ``adjoint_workspace_bytes()``."""
        lib = _lib.load()
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        if workspace is None:
            workspace = torch.empty(self.adjoint_workspace_bytes(), dtype=torch.uint8,
                                    device=self.t["y0"].device)
        T = self.t
        io = _lib.AdjointIO(ckpt=T["ckpt"].data_ptr(), stats=T["stats"].data_ptr(),
                            ys_bar=ys_bar.data_ptr(), y0_bar=y0_bar.data_ptr(),
                            params_bar=params_bar.data_ptr())
        _lib.check(lib.dfx_erk_adjoint(ctypes.byref(self.problem), ctypes.byref(self.io),
                                       ctypes.byref(io), ctypes.c_void_p(workspace.data_ptr()),
                                       ctypes.c_size_t(workspace.numel()),
                                       ctypes.c_void_p(stream)), "dfx_erk_adjoint")

    def adjoint_workspace_bytes(self):
        nbytes = ctypes.c_size_t(0)
        _lib.check(_lib.load().dfx_erk_adjoint_workspace_bytes(ctypes.byref(self.problem),
                                                                ctypes.byref(nbytes)),
                   "dfx_erk_adjoint_workspace_bytes")
        return int(nbytes.value)

    def launch(self, stream=None):
        lib = _lib.load()
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        lp = None
        if self.launch_cfg is not None:
            lp = ctypes.byref(_lib.Launch(int(self.launch_cfg[0]), int(self.launch_cfg[1])))
        rc = lib.dfx_erk_solve(ctypes.byref(self.problem), ctypes.byref(self.io),
                               ctypes.c_void_p(self.ws.data_ptr()), ctypes.c_size_t(self.ws.numel()),
                               lp, ctypes.c_void_p(stream))
        _lib.check(rc, "dfx_erk_solve")


def _check_ts(ts, t0, t1_vals, directions):
    """saveat.ts must be increasing and lie in [t0, t1] (_integrate.py:1222-1235).
    ts: [n] or [B, n]; t1_vals, directions: [1] or [B] (float64 numpy)."""
    ts = np.asarray(ts, dtype=np.float64)
    dirs = np.asarray(directions, dtype=np.float64).reshape(-1, 1)
    tn = np.atleast_2d(ts) * dirs
    if tn.shape[-1] > 1 and bool((tn[:, 1:] < tn[:, :-1]).any()):
        raise RuntimeError("saveat.ts must be increasing or decreasing.")
    t0n = t0 * dirs
    t1n = np.asarray(t1_vals, dtype=np.float64).reshape(-1, 1) * dirs
    if bool(((tn > t1n) | (tn < t0n)).any()):
        raise RuntimeError("saveat.ts must lie between t0 and t1.")


def _controller_fields(stepsize_controller, dt0f):
    """The dfx_erk_problem fields of a PIDController or ConstantStepSize."""
    if isinstance(stepsize_controller, PIDController):
        c = stepsize_controller
        if c.norm is not rms_norm:
            raise NotImplementedError(
                "adjoint_rms_seminorm is provided as the norm of the neural-ODE "
                "BacksolveAdjoint's controller; the small-system kernels use rms_norm")
        per_instance = np.ndim(c.rtol) > 0 or np.ndim(c.atol) > 0
        return dict(controller=_lib.PID, pcoeff=float(c.pcoeff), icoeff=float(c.icoeff),
                    dcoeff=float(c.dcoeff), rtol=0.0 if per_instance else c.rtol,
                    atol=0.0 if per_instance else c.atol,
                    dtmin=_nan_if_none(c.dtmin), dtmax=_nan_if_none(c.dtmax),
                    factormin=c.factormin, factormax=c.factormax, safety=c.safety,
                    error_order=_nan_if_none(c.error_order), force_dtmin=int(c.force_dtmin))
    if isinstance(stepsize_controller, ConstantStepSize):
        if dt0f is None:
            raise ValueError("Constant step size solvers cannot select step size "
                             "automatically; please pass a value for `dt0`.")
        return dict(controller=_lib.CONSTANT, pcoeff=0.0, icoeff=1.0, dcoeff=0.0, rtol=0.0,
                    atol=0.0, dtmin=float("nan"), dtmax=float("nan"), factormin=0.2,
                    factormax=10.0, safety=0.9, error_order=float("nan"), force_dtmin=1)
    raise NotImplementedError(f"stepsize controller {stepsize_controller!r}")


def _state_record(solver, stepsize_controller, B, d, batched, solver_state, controller_state,
                  made_jump):
    """The resume mask and the [B, d + 9] float64 state records of dfx_erk_io.state_in
    (the record layout of DFX_ERK_STATE_WORDS) from diffeqsolve's solver_state,
    controller_state and made_jump (_integrate.py:906-908, 1252-1270, 1309); each value is
    a scalar shared by the batch or has the batch as its leading dimension."""
    rec = torch.zeros((B, d + 9), dtype=torch.float64)
    mask = 0

    def col(v, name):
        t = torch.as_tensor(v).detach().to("cpu", torch.float64).reshape(-1)
        if t.numel() not in (1, B):
            raise ValueError(f"{name} must be a scalar or have shape [{B}]")
        return t.expand(B)

    if solver_state is not None and solver.tableau.fsal:
        if not (isinstance(solver_state, (tuple, list)) and len(solver_state) == 2):
            raise ValueError("solver_state must be the (first_step, f0) pair of a previous "
                             "solve's SaveAt(solver_state=True)")
        first, f0 = solver_state
        f0 = torch.as_tensor(f0).detach().to("cpu", torch.float64)
        if f0.shape not in ((B, d), (d,)) or (f0.ndim == 1 and batched and B > 1):
            raise ValueError(f"solver_state f0 must have shape [{B}, {d}]" if batched
                             else f"solver_state f0 must have shape [{d}]")
        rec[:, :d] = f0.reshape(-1, d).expand(B, d)
        rec[:, d] = col(first, "solver_state first_step")
        mask |= _lib.RESUME_SOLVER
    if made_jump is not None:
        rec[:, d + 1] = col(made_jump, "made_jump")
        mask |= _lib.RESUME_MADE_JUMP
    if controller_state is not None:
        cs = tuple(controller_state)
        if isinstance(stepsize_controller, PIDController):
            if len(cs) != 3:
                raise ValueError("a PIDController state is (prev_inv_scaled_error, "
                                 "prev_prev_inv_scaled_error, at_dtmin)")
            for i, name in enumerate(("prev_inv", "prev_prev_inv", "at_dtmin")):
                rec[:, d + 2 + i] = col(cs[i], name)
        else:
            if len(cs) != 4:
                raise ValueError("a ConstantStepSize state is (steps_completed, num_steps, "
                                 "t0_sim, t1_sim_or_dt0)")
            for i, name in enumerate(("steps_completed", "num_steps", "t0_sim", "t1_sim")):
                rec[:, d + 5 + i] = col(cs[i], name)
        mask |= _lib.RESUME_CONTROLLER
    return mask, rec


def _state_outputs(plan, saveat):
    """Solution.solver_state / controller_state / made_jump from the final state records
    (_integrate.py:1489-1500), as device tensors with the batch as leading dimension
    (dropped for an unbatched solve).  The PID state is float64 in both precisions."""
    rec = plan.t.get("state_out")
    if rec is None:
        return None, None, None
    d = plan.problem.d
    pick = (lambda v: v) if plan.batched else (lambda v: v[0])
    solver_state = controller_state = made_jump = None
    if saveat.solver_state and plan.fsal:
        solver_state = (pick(rec[:, d] != 0), pick(rec[:, :d].to(plan.t["y0"].dtype)))
    if saveat.controller_state:
        if plan.problem.controller == _lib.PID:
            controller_state = (pick(rec[:, d + 2]), pick(rec[:, d + 3]), pick(rec[:, d + 4] != 0))
        else:
            controller_state = (pick(rec[:, d + 5].to(torch.int32)),
                                pick(rec[:, d + 6].to(torch.int32)), pick(rec[:, d + 7]),
                                pick(rec[:, d + 8]))
    if saveat.made_jump:
        made_jump = pick(rec[:, d + 1] != 0)
    return solver_state, controller_state, made_jump


def _prepare_erk(terms, solver, t0, t1, dt0, y0, args, saveat, stepsize_controller,
                 max_steps, launch, record_grid=False, event=None, solver_state=None,
                 controller_state=None, made_jump=None):
    # Host-side checks first (they need no GPU); device placement at the end.
    if not isinstance(terms, ODETerm) or not isinstance(terms.vector_field, RegisteredVectorField):
        raise NotImplementedError(
            "diffrax_amd evaluates registered device vector fields only: wrap one of "
            "LotkaVolterra, VanDerPol, LinearField, MLPField in ODETerm.")
    vfo = terms.vector_field
    if not isinstance(solver, AbstractERK):
        raise NotImplementedError(f"solver {solver!r} is not an explicit Runge-Kutta method")
    y0 = torch.as_tensor(y0)
    if y0.dtype not in (torch.float64, torch.float32):
        raise ValueError(f"y0 must be float64 or float32, got {y0.dtype}")
    batched = y0.ndim == 2
    y0b = y0 if batched else y0.reshape(1, -1)
    B, d = y0b.shape
    if not vfo.supports(d):
        raise NotImplementedError(f"{type(vfo).__name__} does not support state dimension {d}")
    precision = _lib.F64 if y0b.dtype == torch.float64 else _lib.F32S_F64T

    t0f = _scalar_time(t0, "t0")
    if t0f is None:
        raise ValueError("t0 must be a scalar")
    host = _HostCopies()
    if isinstance(t1, torch.Tensor) and t1.device.type != "cpu":
        # device-resident t1 (one value or one per instance): read by the kernel in place
        t1_t = t1.to(torch.float64).reshape(-1)
        if t1_t.numel() not in (1, B):
            raise ValueError(f"t1 must be a scalar or have shape [{B}]")
        t1_stride = 0 if t1_t.numel() == 1 else 1
        host.add("t1", t1_t)
    else:
        t1f = _scalar_time(t1, "t1")
        if t1f is not None:
            t1_t = torch.tensor([t1f], dtype=torch.float64)
            t1_stride = 0
        else:
            t1_t = torch.as_tensor(t1, dtype=torch.float64).reshape(-1)
            if t1_t.numel() != B:
                raise ValueError(f"t1 must be a scalar or have shape [{B}]")
            t1_stride = 1
        host.add("t1", t1_t)

    if dt0 is not None:
        dt0f = _scalar_time(dt0, "dt0")
        if dt0f is None:
            raise NotImplementedError("dt0 must be a scalar")
    else:
        dt0f = None

    want_state = saveat.solver_state or saveat.controller_state or saveat.made_jump
    resuming = solver_state is not None or controller_state is not None or made_jump is not None
    if (want_state or resuming) and isinstance(stepsize_controller, ClipStepSizeController):
        raise NotImplementedError("solver/controller state export and resume are provided "
                                  "for PIDController and ConstantStepSize")
    sub = saveat.subs
    if sub is None:
        if not saveat.dense and not want_state:
            raise ValueError("SaveAt must save something")
        sub = SubSaveAt.__new__(SubSaveAt)   # dense output / state only: no saved rows
        sub.t0 = sub.t1 = False
        sub.ts, sub.steps, sub.fn = None, 0, save_y
    if not hasattr(sub, "fn"):
        raise NotImplementedError("only a single SubSaveAt is supported")
    if sub.steps < 0:
        raise ValueError("SaveAt(steps=n) needs n >= 0")
    if max_steps is None:
        if saveat.dense:
            raise ValueError("`max_steps=None` is incompatible with `saveat.dense=True`")
        raise NotImplementedError("max_steps=None is not supported; pass an integer")

    ts_t = None
    ts_stride = 0
    n_ts = 0
    if sub.ts is not None:
        ts_t = torch.as_tensor(sub.ts, dtype=torch.float64)
        if ts_t.ndim == 1:
            n_ts = ts_t.shape[0]
        elif ts_t.ndim == 2 and ts_t.shape[0] == B:
            n_ts = ts_t.shape[1]
            ts_stride = n_ts
        else:
            raise ValueError(f"SaveAt(ts) must be [n] or [{B}, n]")
        host.add("ts", ts_t)

    def check_times():
        # _integrate.py:1107-1112 (dt0 sign) and 1222-1235 (ts order and range)
        t1_host = host.get("t1")
        if dt0f is not None and bool(np.any((t1_host - t0f) * dt0f < 0)):
            raise RuntimeError(
                f"Must have (t1 - t0) * dt0 >= 0, we instead got t1 with value {t1}, t0 "
                f"with value {t0}, dt0 with value {dt0}")
        if ts_t is not None:
            _check_ts(host.get("ts"), t0f, t1_host, np.where(t0f < t1_host, 1.0, -1.0))

    host.seal()
    if not host.deferred:
        check_times()

    clip = None
    if isinstance(stepsize_controller, ClipStepSizeController):
        clip = stepsize_controller
        if not isinstance(clip.controller, PIDController):
            raise NotImplementedError(
                "ClipStepSizeController is provided around a PIDController only")
        if clip.store_rejected_steps is not None or clip.callback_on_reject is not None:
            raise NotImplementedError("store_rejected_steps is not provided by diffrax_amd")
        stepsize_controller = clip.controller
    ctrl = _controller_fields(stepsize_controller, dt0f)
    if resuming:
        resume_mask, state_rec = _state_record(solver, stepsize_controller, B, d, batched,
                                               solver_state, controller_state, made_jump)

    device = _require_gpu()
    y0b = y0b.to(device).contiguous()
    t1_t = t1_t.to(device).contiguous()
    if ts_t is not None:
        ts_t = ts_t.to(device).contiguous()
    params, param_stride = vfo.pack(args, B, d, y0b.dtype, device)
    prob = _lib.ErkProblem(
        solver=solver.solver_id, vf=vfo.vf_id, precision=precision,
        d=d, hidden=0, n_ts=n_ts, save_t0=int(sub.t0), save_t1=int(sub.t1),
        save_steps=int(sub.steps), max_steps=int(max_steps), batch=B, t1_stride=t1_stride,
        ts_stride=ts_stride, param_stride=param_stride, t0=t0f,
        dt0=_nan_if_none(dt0f), **ctrl)
    ev = {}
    if event is not None:
        prob.n_events = len(event.conds)
        prob.event_root_find = int(event.root_finder is not None)
        ev["event_coeffs"] = torch.tensor([c.coeffs(d) for c in event.conds],
                                          dtype=torch.float64, device=device)
        ev["event_flags"] = torch.tensor(event.flags(), dtype=torch.int32, device=device)
        ev["event_mask"] = torch.zeros((B, len(event.conds)), dtype=torch.int32, device=device)
    clip_t = {}
    for name in ("step_ts", "jump_ts"):
        arr = None if clip is None else getattr(clip, name)
        if arr is not None and arr.size > 0:
            clip_t[name] = torch.tensor(arr, dtype=torch.float64, device=device)
            setattr(prob, "n_" + name, int(arr.size))
    tol_t = {}
    if isinstance(stepsize_controller, PIDController) and (
            np.ndim(stepsize_controller.rtol) > 0 or np.ndim(stepsize_controller.atol) > 0):
        for name in ("rtol", "atol"):
            v = np.asarray(getattr(stepsize_controller, name), dtype=np.float64)
            if v.ndim > 0 and v.shape[0] != B:
                raise ValueError(f"per-instance {name} must have shape [{B}]")
            tol_t[name] = torch.tensor(np.broadcast_to(v, (B,)).copy(), dtype=torch.float64,
                                       device=device)
    lib = _lib.load()
    n_out = lib.dfx_erk_n_out(ctypes.byref(prob))
    tensors = dict(
        y0=y0b, t1=t1_t, ts=ts_t, params=params,
        ys=torch.empty((B, n_out, d), dtype=y0b.dtype, device=device),
        ts_out=torch.empty((B, n_out), dtype=torch.float64, device=device),
        stats=torch.empty((B, 3), dtype=torch.int32, device=device),
        result=torch.empty((B,), dtype=torch.int32, device=device),
        grid=(torch.empty((B, int(max_steps), 2), dtype=torch.float64, device=device)
              if record_grid else None),
        step_ts=clip_t.get("step_ts"), jump_ts=clip_t.get("jump_ts"), **ev, **tol_t,
    )
    if resuming:
        prob.resume = resume_mask
        tensors["state_in"] = state_rec.to(device)
    if want_state:
        tensors["state_out"] = torch.empty((B, d + 9), dtype=torch.float64, device=device)
    if saveat.dense:
        S = solver.tableau.num_stages
        tensors.update(
            dense_ts=torch.empty((B, int(max_steps) + 1), dtype=torch.float64, device=device),
            dense_ys=torch.empty((B, int(max_steps) + 1, d), dtype=y0b.dtype, device=device),
            dense_k=torch.empty((B, int(max_steps), S, d), dtype=y0b.dtype, device=device))
    plan = ErkPlan(prob, tensors, n_out, batched, None if launch == "auto" else launch)
    plan.fsal = bool(solver.tableau.fsal)
    # checks of device-resident times, run by diffeqsolve once the solve is queued
    plan.deferred_checks = check_times if host.deferred else None
    return plan


def _finish(plan, t0, t1, max_steps, throw, ys=None, saveat=None):
    if plan.deferred_checks is not None:
        plan.deferred_checks()
    T = plan.t
    stats = {
        "num_steps": T["stats"][:, 0],
        "num_accepted_steps": T["stats"][:, 1],
        "num_rejected_steps": T["stats"][:, 2],
        "max_steps": max_steps,
    }
    ys = T["ys"] if ys is None else ys
    ts, result = T["ts_out"], T["result"]
    if not plan.batched:
        ys, ts, result = ys[0], ts[0], result[0]
        stats = {k: (v[0] if isinstance(v, torch.Tensor) else v) for k, v in stats.items()}
    if throw:
        codes = T["result"]
        bad = (codes != 0) & (codes != int(RESULTS.event_occurred))   # is_okay
        if bool(bad.any()):
            first = int(codes[bad][0].item())
            raise RuntimeError(RESULTS[first])
    res = result
    interp = DenseInterpolation(plan) if T.get("dense_ts") is not None else None
    if plan.problem.n_ts == 0 and not (plan.problem.save_t0 or plan.problem.save_t1 or
                                       plan.problem.save_steps):
        ts = ys = None   # SaveAt(dense=True) alone saves no rows (as the reference)
    event_mask = None
    if T.get("event_mask") is not None:
        m = T["event_mask"].bool()
        if not plan.batched:
            m = m[0]
        ev = getattr(plan, "event", None)
        event_mask = [m[..., i] for i in range(m.shape[-1])] if ev is not None and ev.is_list \
            else m[..., 0]
    solver_state, controller_state, made_jump = (
        _state_outputs(plan, saveat) if saveat is not None else (None, None, None))
    sol = Solution(t0=t0, t1=t1, ts=ts, ys=ys, stats=stats, result=res,
                   interpolation=interp, event_mask=event_mask, solver_state=solver_state,
                   controller_state=controller_state, made_jump=made_jump)
    return sol


class DenseInterpolation:
    """Dense output of a SaveAt(dense=True) solve (DenseInterpolation,
    _global_interpolation.py:315-400), evaluated on the GPU by dfx_erk_dense_eval.

    ``evaluate(t0, t1=None, left=True)`` and ``derivative(t, left=True)`` take a scalar
    time, a 1-D tensor of times shared by the batch, or a [B, n] tensor of per-instance
    times; they return [..., d] for every instance (the batch axis first when the solve was
    batched, then the query axis when t was not a scalar)."""

    def __init__(self, plan):
        self.plan = plan
        T = plan.t
        self.t0 = plan.problem.t0
        self.t1 = T["t1"]

    def _query(self, t, derivative, left):
        lib = _lib.load()
        plan = self.plan
        T = plan.t
        dev = T["y0"].device
        B, d = T["y0"].shape
        tq = torch.as_tensor(t, dtype=torch.float64)
        scalar = tq.ndim == 0
        if tq.ndim == 0:
            tq = tq.reshape(1)
        if tq.ndim == 1:
            n_q, stride = tq.shape[0], 0
        elif tq.ndim == 2 and tq.shape[0] == B:
            n_q, stride = tq.shape[1], tq.shape[1]
        else:
            raise ValueError(f"evaluation times must be a scalar, [n] or [{B}, n]")
        tq = tq.to(dev).contiguous()
        out = torch.empty((B, n_q, d), dtype=T["y0"].dtype, device=dev)
        io = _lib.DenseIO(dense_ts=T["dense_ts"].data_ptr(), dense_ys=T["dense_ys"].data_ptr(),
                          dense_k=T["dense_k"].data_ptr(), stats=T["stats"].data_ptr(),
                          t1=T["t1"].data_ptr(), tq=tq.data_ptr(), n_q=n_q, tq_stride=stride,
                          left=int(bool(left)), derivative=int(bool(derivative)),
                          out=out.data_ptr())
        _lib.check(lib.dfx_erk_dense_eval(ctypes.byref(plan.problem), ctypes.byref(io),
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "dfx_erk_dense_eval")
        if scalar:
            out = out[:, 0]
        if not plan.batched:
            out = out[0]
        return out

    def evaluate(self, t0, t1=None, left: bool = True):
        if t1 is not None:
            return self.evaluate(t1, left=left) - self.evaluate(t0, left=left)
        return self._query(t0, False, left)

    def derivative(self, t, left: bool = True):
        return self._query(t, True, left)


def diffeqsolve(terms, solver: AbstractSolver, t0, t1, dt0, y0, args: Any = None, *,
                saveat: SaveAt = SaveAt(t1=True), stepsize_controller=ConstantStepSize(),
                adjoint: AbstractAdjoint = RecursiveCheckpointAdjoint(), event=None,
                max_steps: Optional[int] = 4096, throw: bool = True, progress_meter=None,
                solver_state=None, controller_state=None, made_jump=None,
                launch: Optional[tuple] = None) -> Solution:
    """Solve an initial value problem on the GPU (signature of _integrate.py:890-911).

    Extra keyword ``launch``: ``(block, lanes)`` sets the small-system kernel's launch
    shape, 4 or 8 the neural-ODE kernel's wavefronts per tile; ``"auto"`` times the
    candidates once per problem signature and device and keeps the fastest
    (``_autotune``).
    Batched use: y0 of shape [B, d] solves B independent instances in one launch.

    SaveAt(fn=...) is applied to the saved (t, y) rows after the solve, batched with
    torch.func.vmap (unfilled rows stay inf, as the reference presets them,
    _integrate.py:1272-1305); SaveAt(subs=[...]) with several SubSaveAts runs one solve per
    SubSaveAt and returns ts / ys in the same container (the solves are deterministic, so
    each equals what one combined solve would save)."""
    kw = dict(stepsize_controller=stepsize_controller, adjoint=adjoint, event=event,
              max_steps=max_steps, throw=throw, progress_meter=progress_meter,
              solver_state=solver_state, controller_state=controller_state,
              made_jump=made_jump, launch=launch)
    if isinstance(adjoint, BacksolveAdjoint):
        _check_backsolve(saveat, event, adjoint, stepsize_controller)
    subs = saveat.subs
    if subs is not None and not isinstance(subs, SubSaveAt):
        leaves, spec = torch.utils._pytree.tree_flatten(subs)
        if not leaves or not all(isinstance(x, SubSaveAt) for x in leaves):
            raise ValueError("SaveAt(subs=...) must be a SubSaveAt or a container of them")
        first_only = lambda i: dict(dense=saveat.dense and i == 0,
                                    solver_state=saveat.solver_state and i == 0,
                                    controller_state=saveat.controller_state and i == 0,
                                    made_jump=saveat.made_jump and i == 0)
        sols = [diffeqsolve(terms, solver, t0, t1, dt0, y0, args,
                            saveat=SaveAt(subs=leaf, **first_only(i)), **kw)
                for i, leaf in enumerate(leaves)]
        first = sols[0]
        first.ts = torch.utils._pytree.tree_unflatten([x.ts for x in sols], spec)
        first.ys = torch.utils._pytree.tree_unflatten([x.ys for x in sols], spec)
        return first
    fn = save_y if subs is None else subs.fn
    if fn is not save_y:
        plain = SubSaveAt.__new__(SubSaveAt)
        plain.__dict__.update(subs.__dict__)
        plain.fn = save_y
        inner = SaveAt.__new__(SaveAt)
        inner.__dict__.update(saveat.__dict__)
        inner.subs = plain
        sol = _diffeqsolve_one(terms, solver, t0, t1, dt0, y0, args, saveat=inner, **kw)
        sol.ys = _apply_save_fn(fn, sol.ts, sol.ys, args, torch.as_tensor(y0).ndim == 2)
        return sol
    return _diffeqsolve_one(terms, solver, t0, t1, dt0, y0, args, saveat=saveat, **kw)


def _apply_save_fn(fn, ts, ys, args, batched):
    """ys rows replaced by fn(t, y, args) (_integrate.py:244-275, 456-527 save fn(t, y,
    args) in the caller's time); rows the solve did not fill stay inf."""
    from torch.func import vmap
    from torch.utils import _pytree as pytree
    per_row = vmap(fn, in_dims=(0, 0, None))
    args = pytree.tree_map(
        lambda a: a.to(ys.device) if isinstance(a, torch.Tensor) else a, args)
    if batched:
        B = ys.shape[0]
        if ts.ndim == 1:
            ts = ts.expand(B, ts.shape[0])
        dims = pytree.tree_map(
            lambda a: 0 if isinstance(a, torch.Tensor) and a.ndim >= 1 and a.shape[0] == B
            else None, args)
        out = vmap(per_row, in_dims=(0, 0, dims))(ts, ys, args)
    else:
        out = per_row(ts, ys, args)
    filled = torch.isfinite(ts)

    def mask(o):
        m = filled.reshape(filled.shape + (1,) * (o.ndim - filled.ndim))
        return torch.where(m.to(o.device), o, torch.full_like(o, float("inf")))

    return pytree.tree_map(mask, out)


def _diffeqsolve_one(terms, solver, t0, t1, dt0, y0, args, *, saveat, stepsize_controller,
                     adjoint, event, max_steps, throw, progress_meter, solver_state,
                     controller_state, made_jump, launch):
    resuming = solver_state is not None or controller_state is not None or made_jump is not None
    if (resuming or saveat.solver_state or saveat.controller_state or saveat.made_jump) and (
            _is_ou_sde(terms)):
        raise NotImplementedError("solver/controller state export and resume are provided "
                                  "for the ODE kernels only")
    if progress_meter is not None and type(progress_meter).__name__ != "NoProgressMeter":
        raise NotImplementedError("progress meters are not provided by diffrax_amd")
    if event is not None and _is_ou_sde(terms):
        raise NotImplementedError("events are provided for the ODE kernels only")
    if _is_ou_sde(terms):
        from ._sde import solve_sde
        device = _require_gpu()
        ts, ys, n, interp = solve_sde(terms, solver, t0, t1, dt0, y0, args, saveat,
                                      stepsize_controller, adjoint, device)
        if n > max_steps:
            if throw:
                raise RuntimeError(RESULTS[RESULTS.max_steps_reached])
        shape = interp.run.ys.shape[:1] if ys is None else ys.shape[:-2]
        if ys is None and not interp.batched:
            shape = ()
        full = lambda v: torch.full(shape, v, dtype=torch.int32, device=device)
        stats = {"num_steps": full(n), "num_accepted_steps": full(n),
                 "num_rejected_steps": full(0), "max_steps": max_steps}
        return Solution(t0=t0, t1=t1, ts=ts, ys=ys, stats=stats,
                        result=full(0), interpolation=interp)
    if isinstance(terms, ODETerm) and isinstance(terms.vector_field, MLPField):
        from . import _mlp
        return _mlp.diffeqsolve_mlp(terms, solver, t0, t1, dt0, y0, saveat,
                                    stepsize_controller, adjoint, max_steps, throw,
                                    _require_gpu(), launch=launch, event=event,
                                    solver_state=solver_state,
                                    controller_state=controller_state, made_jump=made_jump)
    plan = _prepare_erk(terms, solver, t0, t1, dt0, y0, args, saveat, stepsize_controller,
                        max_steps, launch, event=event, solver_state=solver_state,
                        controller_state=controller_state, made_jump=made_jump)
    plan.event = event
    T = plan.t
    wants_grad = torch.is_grad_enabled() and (T["y0"].requires_grad or
                                              T["params"].requires_grad)
    if launch == "auto":
        from . import _autotune
        _autotune.tune(plan)
    if not wants_grad:
        plan.launch()
        return _finish(plan, t0, t1, max_steps, throw, saveat=saveat)
    if resuming or T.get("state_out") is not None:
        raise NotImplementedError("gradients through a solve that resumes or saves "
                                  "solver/controller state are not provided")
    if isinstance(adjoint, BacksolveAdjoint):
        back = _backsolve_problem(plan.problem, adjoint, stepsize_controller)
        ys = _ErkBacksolve.apply(plan, back, bool(adjoint.kwargs.get("throw", throw)),
                                 T["y0"], T["params"])
        return _finish(plan, t0, t1, max_steps, throw, ys=ys)
    if not isinstance(adjoint, RecursiveCheckpointAdjoint):
        raise NotImplementedError(f"adjoint {adjoint!r} is not provided by diffrax_amd")
    _check_discrete_adjoint(plan, event, saveat)
    plan.record_checkpoints(adjoint.checkpoints)
    ys = _ErkDiscreteAdjoint.apply(plan, T["y0"], T["params"])
    return _finish(plan, t0, t1, max_steps, throw, ys=ys)


def _check_discrete_adjoint(plan, event, saveat):
    p = plan.problem
    if event is not None:
        raise NotImplementedError("gradients through a solve with events are not provided")
    if saveat.dense:
        raise NotImplementedError(
            "gradients through SaveAt(dense=True) are not provided: Solution.evaluate is not "
            "differentiable here")
    ok = (p.vf in (_lib.VF_LV, _lib.VF_VDP) and p.d == 2) or (p.vf == _lib.VF_LINEAR and
                                                               1 <= p.d <= 4)
    if not ok:
        raise NotImplementedError("gradients of the small-system solve are provided for "
                                  "LotkaVolterra, VanDerPol and LinearField with d <= 4")


def _check_backsolve(saveat, event, adjoint, stepsize_controller):
    """The guards of BacksolveAdjoint.loop (_adjoint.py:798-858), and the backward
    controller this build provides, checked before any GPU work."""
    if saveat.subs is not None and not hasattr(saveat.subs, "fn"):
        raise NotImplementedError(
            "Cannot use `adjoint=BacksolveAdjoint()` with `SaveAt(subs=...)`.")
    if saveat.dense or (saveat.subs is not None and saveat.subs.steps):
        raise NotImplementedError(
            "Cannot use `adjoint=BacksolveAdjoint()` with `saveat=SaveAt(steps=True)` or "
            "saveat=SaveAt(dense=True)`.")
    if saveat.subs is not None and saveat.subs.fn is not save_y:
        raise NotImplementedError(
            "Cannot use `adjoint=BacksolveAdjoint()` with `saveat=SaveAt(fn=...)`.")
    if event is not None:
        raise NotImplementedError("`diffrax.BacksolveAdjoint` is not compatible with events.")
    ctrl = adjoint.kwargs.get("stepsize_controller", stepsize_controller)
    if isinstance(ctrl, ClipStepSizeController):
        raise NotImplementedError("BacksolveAdjoint with ClipStepSizeController is not provided")
    if isinstance(ctrl, PIDController) and (np.ndim(ctrl.rtol) > 0 or np.ndim(ctrl.atol) > 0):
        raise NotImplementedError("BacksolveAdjoint with per-instance tolerances is not provided")
    if isinstance(ctrl, PIDController) and (float(ctrl.pcoeff) != 0.0 or
                                            float(ctrl.dcoeff) != 0.0 or ctrl.dtmin is not None):
        raise NotImplementedError(
            "BacksolveAdjoint is provided with integral step control (pcoeff = dcoeff = 0) "
            "and no dtmin: the backward solves then carry no controller state between save "
            "times")
    if "solver" in adjoint.kwargs and not isinstance(adjoint.kwargs["solver"], AbstractERK):
        raise NotImplementedError(
            f"backward solver {adjoint.kwargs['solver']!r} is not an explicit Runge-Kutta method")
    if adjoint.kwargs.get("max_steps", 0) is None:
        raise NotImplementedError("max_steps=None is not supported; pass an integer")


def _backsolve_problem(prob, adjoint, stepsize_controller):
    """The problem the backward solves run: the forward's, with BacksolveAdjoint's kwargs
    applied as the reference applies them (_adjoint.py:623-632).  dfx_erk_backsolve solves
    with -problem.dt0, so a dt0 given to BacksolveAdjoint is stored negated."""
    kw = adjoint.kwargs
    back = _lib.ErkProblem()
    ctypes.memmove(ctypes.byref(back), ctypes.byref(prob), ctypes.sizeof(prob))
    if "solver" in kw:
        back.solver = kw["solver"].solver_id
    if "dt0" in kw:
        d = kw["dt0"]
        back.dt0 = float("nan") if d is None else -float(d)
    dt0b = None if math.isnan(back.dt0) else -back.dt0
    ctrl = kw.get("stepsize_controller", stepsize_controller)
    for k, v in _controller_fields(ctrl, dt0b).items():
        setattr(back, k, v)
    if "max_steps" in kw:
        back.max_steps = int(kw["max_steps"])
    return back


class _ErkDiscreteAdjoint(torch.autograd.Function):
    """ys of a small-system solve, with reverse mode through the accepted steps
    (dfx_erk_adjoint, replaying each segment from the forward's checkpoints) as its
    backward."""

    @staticmethod
    def forward(ctx, plan, y0, params):
        plan.launch()
        ctx.plan = plan
        return plan.t["ys"]

    @staticmethod
    def backward(ctx, gys):
        plan = ctx.plan
        T = plan.t
        B, d = T["y0"].shape
        ys = T["ys"]
        gys = torch.zeros_like(ys) if gys is None else gys.to(ys.dtype).contiguous()
        P = T["params"].shape[-1]
        y0_bar = torch.empty((B, d), dtype=ys.dtype, device=ys.device)
        params_bar = torch.empty((B, P), dtype=ys.dtype, device=ys.device)
        plan.adjoint(gys, y0_bar, params_bar)
        gparams = params_bar.sum(0) if plan.problem.param_stride == 0 else params_bar
        return None, y0_bar, gparams.reshape(T["params"].shape)


class _ErkBacksolve(torch.autograd.Function):
    """ys of a small-system solve, with the continuous adjoint as its backward.  The saved
    rows are the backward solves' restart states, so they are kept through
    save_for_backward: an in-place change of the returned ys makes backward raise."""

    @staticmethod
    def forward(ctx, plan, back, throw, y0, params):
        plan.launch()
        ctx.plan, ctx.back, ctx.throw = plan, back, throw
        ys = plan.t["ys"]
        ctx.save_for_backward(ys)
        return ys

    @staticmethod
    def backward(ctx, gys):
        plan, back = ctx.plan, ctx.back
        T = plan.t
        B, d = T["y0"].shape
        ys, = ctx.saved_tensors
        gys = torch.zeros_like(ys) if gys is None else gys.to(ys.dtype).contiguous()
        P = T["params"].shape[-1]
        y0_bar = torch.empty((B, d), dtype=ys.dtype, device=ys.device)
        params_bar = torch.empty((B, P), dtype=ys.dtype, device=ys.device)
        result = torch.empty((B,), dtype=torch.int32, device=ys.device)
        io = _lib.BacksolveIO(params=T["params"].data_ptr(), ys=ys.data_ptr(),
                              ts_out=T["ts_out"].data_ptr(), ys_bar=gys.data_ptr(),
                              y0_bar=y0_bar.data_ptr(), params_bar=params_bar.data_ptr(),
                              result=result.data_ptr())
        lib = _lib.load()
        _lib.check(lib.dfx_erk_backsolve(ctypes.byref(back), ctypes.byref(io),
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "dfx_erk_backsolve")
        if ctx.throw and bool((result != 0).any()):
            raise RuntimeError(RESULTS[int(result[result != 0][0].item())])
        gparams = params_bar.sum(0) if plan.problem.param_stride == 0 else params_bar
        return None, None, None, y0_bar, gparams.reshape(T["params"].shape)


def _is_ou_sde(terms):
    if not isinstance(terms, MultiTerm) or len(terms.terms) != 2:
        return False
    a, b = terms.terms
    return (isinstance(a, ODETerm) and isinstance(a.vector_field, OUDrift)
            and isinstance(b, ControlTerm) and isinstance(b.vector_field, DiagonalNoise)
            and isinstance(b.control, IncrementPath))


_ = math
