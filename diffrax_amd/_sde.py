"""Heun SDE path: MultiTerm(ODETerm(OUDrift()), ControlTerm(DiagonalNoise(),
IncrementPath(dW))) with Heun and ConstantStepSize, forward and discrete adjoint on the
GPU (sde_heun.hip)."""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib


class SdeProblem(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("d", ctypes.c_int32),
                ("paths", ctypes.c_int64), ("num_steps", ctypes.c_int32),
                ("checkpoint_every", ctypes.c_int32), ("save_t0", ctypes.c_int32),
                ("save_t1", ctypes.c_int32), ("save_steps", ctypes.c_int32),
                ("n_ts", ctypes.c_int32)]


class SdeIO(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("y0", "theta", "mu", "sigma", "dW", "grid", "ys", "ckpt", "ts", "dense_ys",
                 "dense_k")]


class SdeDenseIO(ctypes.Structure):
    _fields_ = [("dense_ys", ctypes.c_void_p), ("dense_k", ctypes.c_void_p),
                ("grid", ctypes.c_void_p), ("tq", ctypes.c_void_p), ("n_q", ctypes.c_int64),
                ("tq_stride", ctypes.c_int64), ("left", ctypes.c_int32),
                ("derivative", ctypes.c_int32), ("out", ctypes.c_void_p)]


class SdeGradIO(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("ys_bar", "y0_bar", "theta_bar", "mu_bar", "sigma_bar")]


_SIGS = {
    "dfx_sde_n_out": ([ctypes.POINTER(SdeProblem)], ctypes.c_int),
    "dfx_sde_n_ckpt": ([ctypes.POINTER(SdeProblem)], ctypes.c_int),
    "dfx_sde_workspace_bytes": ([ctypes.POINTER(SdeProblem), ctypes.POINTER(ctypes.c_size_t)],
                                ctypes.c_int),
    "dfx_sde_heun_solve": ([ctypes.POINTER(SdeProblem), ctypes.POINTER(SdeIO), ctypes.c_void_p],
                           ctypes.c_int),
    "dfx_sde_heun_adjoint": ([ctypes.POINTER(SdeProblem), ctypes.POINTER(SdeIO),
                              ctypes.POINTER(SdeGradIO), ctypes.c_void_p, ctypes.c_size_t,
                              ctypes.c_void_p], ctypes.c_int),
    "dfx_sde_heun_adjoint_static": ([ctypes.POINTER(SdeProblem), ctypes.POINTER(SdeIO),
                                     ctypes.POINTER(SdeGradIO), ctypes.c_void_p], ctypes.c_int),
    "dfx_sde_dense_eval": ([ctypes.POINTER(SdeProblem), ctypes.POINTER(SdeDenseIO),
                            ctypes.c_void_p], ctypes.c_int),
}
# DFX_SDE_STATIC_SCRATCH_BYTES: the static partial-sum array of dfx_sde_heun_adjoint_static
SDE_STATIC_SCRATCH_BYTES = 8 << 20
_lib._SIGS.update(_SIGS)


def constant_grid(t0: float, t1: float, dt0: float):
    """ConstantStepSize step times (constant.py:30-104) with the loop's end clipping
    (_integrate.py:278-284, 319-322).  Forward time only."""
    num = int(math.ceil((t1 - t0) / float(np.nextafter(dt0, math.inf))))
    floor = t1
    for _ in range(100):
        floor = float(np.nextafter(floor, -math.inf))
    ts = [t0]
    tnext = min(t0 + dt0, t1)
    completed = 1
    tprev = t0
    while tprev < t1:
        ts.append(tnext)
        completed += 1
        tn = t0 + (t1 - t0) * (completed / num)
        if completed == num:
            tn = t1
        tprev = min(tnext, t1)
        tnext = t1 if tn > floor else tn
    return np.array(ts, dtype=np.float64)


def _ck_for(num_steps, checkpoints):
    if checkpoints is not None:
        want = max(1, math.ceil(num_steps / max(1, int(checkpoints))))
    else:
        want = max(1, round(math.sqrt(num_steps)))
    for ck in (8, 16, 32):
        if want <= ck:
            return ck
    return 32


class _Sde:
    def __init__(self, y0, theta, mu, sigma, dW, grid, save_t0, save_t1, need_grad, checkpoints,
                 save_steps=0, ts=None, dense=False):
        lib = _lib.load()
        self.P, self.d = y0.shape
        self.N = grid.numel() - 1
        prec = _lib.F64 if y0.dtype == torch.float64 else _lib.F32
        ck = _ck_for(self.N, checkpoints) if need_grad else 0
        self.prob = SdeProblem(precision=prec, d=self.d, paths=self.P, num_steps=self.N,
                               checkpoint_every=ck, save_t0=int(save_t0), save_t1=int(save_t1),
                               save_steps=int(save_steps),
                               n_ts=0 if ts is None else int(ts.numel()))
        n_ck = lib.dfx_sde_n_ckpt(ctypes.byref(self.prob)) if need_grad else 0
        self.n_out = int(lib.dfx_sde_n_out(ctypes.byref(self.prob)))
        self.ys = torch.empty((self.P, self.n_out, self.d), dtype=y0.dtype, device=y0.device)
        self.ckpt = (torch.empty((max(n_ck, 1), self.P, self.d), dtype=y0.dtype, device=y0.device)
                     if need_grad else None)
        self.keep = (y0, theta, mu, sigma, dW, grid)
        io = SdeIO()
        for n, v in zip(("y0", "theta", "mu", "sigma", "dW", "grid"), self.keep):
            setattr(io, n, v.data_ptr())
        io.ys = self.ys.data_ptr()
        io.ckpt = self.ckpt.data_ptr() if self.ckpt is not None else None
        self.ts = ts
        io.ts = ts.data_ptr() if ts is not None and ts.numel() else None
        self.dense = None
        if dense:   # SaveAt(dense=True): step-end states and stage increments
            self.dense = dict(
                dense_ys=torch.empty((self.P, self.N + 1, self.d), dtype=y0.dtype, device=y0.device),
                dense_k=torch.empty((self.P, self.N, 2, self.d), dtype=y0.dtype, device=y0.device))
            io.dense_ys = self.dense["dense_ys"].data_ptr()
            io.dense_k = self.dense["dense_k"].data_ptr()
        self.io = io

    def forward(self):
        lib = _lib.load()
        _lib.check(lib.dfx_sde_heun_solve(ctypes.byref(self.prob), ctypes.byref(self.io),
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "dfx_sde_heun_solve")

    def adjoint(self, ys_bar, need_y0):
        lib = _lib.load()
        y0 = self.keep[0]
        dev = y0.device
        ys_bar = ys_bar.to(dtype=y0.dtype).contiguous()
        y0_bar = torch.empty_like(y0) if need_y0 else None
        thb = torch.empty(self.d, dtype=torch.float64, device=dev)
        mub = torch.empty_like(thb)
        sgb = torch.empty_like(thb)
        nbytes = ctypes.c_size_t(0)
        _lib.check(lib.dfx_sde_workspace_bytes(ctypes.byref(self.prob), ctypes.byref(nbytes)),
                   "dfx_sde_workspace_bytes")
        ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev)
        g = SdeGradIO(ys_bar=ys_bar.data_ptr(),
                      y0_bar=y0_bar.data_ptr() if y0_bar is not None else None,
                      theta_bar=thb.data_ptr(), mu_bar=mub.data_ptr(), sigma_bar=sgb.data_ptr())
        _lib.check(lib.dfx_sde_heun_adjoint(ctypes.byref(self.prob), ctypes.byref(self.io),
                                            ctypes.byref(g), ctypes.c_void_p(ws.data_ptr()),
                                            ctypes.c_size_t(ws.numel()),
                                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "dfx_sde_heun_adjoint")
        return y0_bar, thb, mub, sgb


class _SdeFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y0, theta, mu, sigma, dW, grid, save_t0, save_t1, checkpoints,
                save_steps=0, ts=None):
        need = any(ctx.needs_input_grad[:4])
        s = _Sde(y0.detach(), theta.detach(), mu.detach(), sigma.detach(), dW, grid, save_t0,
                 save_t1, need, checkpoints, save_steps, ts)
        s.forward()
        ctx.sde = s
        ctx.param_dtypes = (theta.dtype, mu.dtype, sigma.dtype)
        return s.ys

    @staticmethod
    def backward(ctx, ys_bar):
        s = ctx.sde
        y0_bar, thb, mub, sgb = s.adjoint(ys_bar, ctx.needs_input_grad[0])
        dt, dm, ds = ctx.param_dtypes
        return (y0_bar, thb.to(dt) if ctx.needs_input_grad[1] else None,
                mub.to(dm) if ctx.needs_input_grad[2] else None,
                sgb.to(ds) if ctx.needs_input_grad[3] else None, None, None, None, None, None,
                None, None)


def solve_sde(terms, solver, t0, t1, dt0, y0, args, saveat, stepsize_controller, adjoint,
              device):
    """Heun + ConstantStepSize solve of the OU SDE; returns (ts, ys, num_steps)."""
    from ._solver import Heun
    from ._step_size_controller import ConstantStepSize
    if not isinstance(solver, Heun):
        raise NotImplementedError("the SDE path implements Heun only")
    if not isinstance(stepsize_controller, ConstantStepSize):
        raise NotImplementedError("the SDE path implements ConstantStepSize only")
    sub = saveat.subs
    if sub is None and not saveat.dense:
        raise ValueError("SaveAt must save something")
    if dt0 is None:
        raise ValueError("Constant step size solvers cannot select step size automatically; "
                         "please pass a value for `dt0`.")
    t0f, t1f, dt0f = float(t0), float(t1), float(dt0)
    if not t1f > t0f or not dt0f > 0:
        raise NotImplementedError("the SDE path integrates forward in time with dt0 > 0")
    drift, diff = terms.terms
    dW = diff.control.dW
    y0 = torch.as_tensor(y0)
    batched = y0.ndim == 2
    y0b = (y0 if batched else y0.reshape(1, -1)).to(device).contiguous()
    P, d = y0b.shape
    theta, mu, sigma = (torch.as_tensor(a, dtype=y0b.dtype).to(device) for a in args)
    for name, a in (("theta", theta), ("mu", mu), ("sigma", sigma)):
        if a.shape != (d,):
            raise NotImplementedError(f"{name} must have shape [{d}] (shared by all paths)")
    grid_np = constant_grid(t0f, t1f, dt0f)
    N = len(grid_np) - 1
    dW = torch.as_tensor(dW)
    if not batched and dW.ndim == 2:
        dW = dW.reshape(dW.shape[0], 1, d)
    if dW.shape != (N, P, d):
        raise ValueError(f"dW must have shape [{N}, {P}, {d}], got {tuple(dW.shape)}")
    if dW.dtype != torch.float32:
        raise ValueError("dW must be float32")
    dW = dW.to(device).contiguous()
    grid = torch.tensor(grid_np, dtype=torch.float64, device=device)
    ckpts = adjoint.checkpoints if hasattr(adjoint, "checkpoints") else None
    sub_none = sub is None
    if sub_none:   # dense output only: no saved rows
        from ._saveat import SubSaveAt
        sub = SubSaveAt.__new__(SubSaveAt)
        sub.t0 = sub.t1 = False
        sub.ts, sub.steps = None, 0
    k = int(sub.steps)
    ts_np = None
    ts_dev = None
    if sub.ts is not None:
        ts_np = np.asarray(torch.as_tensor(sub.ts, dtype=torch.float64).cpu().numpy()).reshape(-1)
        if ts_np.size > 1 and (np.diff(ts_np) < 0).any():
            raise RuntimeError("saveat.ts must be increasing (the SDE path runs forward).")
        if ts_np.size and (ts_np[0] < t0f or ts_np[-1] > t1f):
            raise RuntimeError("saveat.ts must lie between t0 and t1.")
        ts_dev = torch.tensor(ts_np, dtype=torch.float64, device=device)
    if saveat.dense:
        if torch.is_grad_enabled() and (y0b.requires_grad or any(
                a.requires_grad for a in (theta, mu, sigma))):
            raise NotImplementedError("gradients through SaveAt(dense=True) are not provided "
                                      "on the SDE kernel")
        run = _Sde(y0b.detach(), theta.detach(), mu.detach(), sigma.detach(), dW, grid,
                   bool(sub.t0) if sub else False, bool(sub.t1) if sub else False, False, None,
                   k, ts_dev, dense=True)
        run.forward()
        ys = run.ys
        interp = SdeDenseInterpolation(run, batched)
    else:
        ys = _SdeFunction.apply(y0b, theta, mu, sigma, dW, grid, bool(sub.t0), bool(sub.t1),
                                ckpts, k, ts_dev)
        interp = None
    # rows: t0; per step, the ts inside it and then its end when k divides the step count;
    # t1 unless the last step was saved (the kernels' order)
    ts_list = [t0f] if sub.t0 else []
    i = 0
    n_ts = 0 if ts_np is None else ts_np.size
    for n in range(N):
        while i < n_ts and ts_np[i] <= grid_np[n + 1]:
            ts_list.append(float(ts_np[i]))
            i += 1
        if k and (n + 1) % k == 0:
            ts_list.append(float(grid_np[n + 1]))
    if sub.t1 and not (k and N % k == 0):
        ts_list.append(t1f)
    ts = torch.tensor(ts_list, dtype=torch.float64, device=device).expand(P, len(ts_list))
    if not batched:
        ys, ts = ys[0], ts[0]
    if sub_none:
        ts = ys = None   # SaveAt(dense=True) alone saves no rows
    return ts, ys, N, interp


class SdeDenseInterpolation:
    """Dense output of an SDE SaveAt(dense=True) solve (DenseInterpolation,
    _global_interpolation.py:315-400, with Heun's Hermite polynomial), evaluated on the GPU
    by dfx_sde_dense_eval.  ``evaluate(t0, t1=None, left=True)`` and
    ``derivative(t, left=True)`` take a scalar time, [n] times shared by the paths or
    [P, n] per-path times."""

    def __init__(self, run, batched):
        self.run, self.batched = run, batched

    def evaluate(self, t0, t1=None, left: bool = True):
        if t1 is not None:
            return self.evaluate(t1, left=left) - self.evaluate(t0, left=left)
        return self._query(t0, False, left)

    def derivative(self, t, left: bool = True):
        return self._query(t, True, left)

    def _query(self, t, derivative, left):
        run = self.run
        dev = run.ys.device
        tq = torch.as_tensor(t, dtype=torch.float64)
        scalar = tq.ndim == 0
        if scalar:
            tq = tq.reshape(1)
        if tq.ndim == 1:
            n_q, stride = tq.shape[0], 0
        elif tq.ndim == 2 and tq.shape[0] == run.P:
            n_q, stride = tq.shape[1], tq.shape[1]
        else:
            raise ValueError(f"evaluation times must be a scalar, [n] or [{run.P}, n]")
        tq = tq.to(dev).contiguous()
        out = torch.empty((run.P, n_q, run.d), dtype=run.ys.dtype, device=dev)
        io = SdeDenseIO(dense_ys=run.dense["dense_ys"].data_ptr(),
                        dense_k=run.dense["dense_k"].data_ptr(), grid=run.keep[5].data_ptr(),
                        tq=tq.data_ptr(), n_q=n_q, tq_stride=stride, left=int(bool(left)),
                        derivative=int(bool(derivative)), out=out.data_ptr())
        _lib.check(_lib.load().dfx_sde_dense_eval(
            ctypes.byref(run.prob), ctypes.byref(io),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "dfx_sde_dense_eval")
        if scalar:
            out = out[:, 0]
        if not self.batched:
            out = out[0]
        return out
