"""CPU oracle for diffrax's explicit Runge-Kutta path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the checker.
The product package ``diffrax_amd`` never imports it.

Two independent restatements live here:

* ``liberk_oracle.so`` (``erk_oracle.c``): FP64 C, batched with OpenMP, every
  registered vector field (Lotka-Volterra, Van der Pol, linear, MLP, OU SDE drift).
  It is the parity reference for the HIP kernels and the timed CPU baseline.
* ``pyoracle.py``: a per-trajectory Python restatement that accepts any Python
  vector field and any array type (numpy, numpy complex, torch with autograd).  It
  pins the C oracle (the two agree on step counts and saved states) and replays a
  fixed accepted-step grid for gradient references.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liberk_oracle.so")

SOLVERS = {"tsit5": 0, "dopri5": 1, "heun": 2}
CONTROLLERS = {"pid": 0, "constant": 1}
VFS = {"lv": 0, "vdp": 1, "linear": 2, "mlp": 3, "ou": 4}
RESULT_NAMES = {0: "successful", 1: "max_steps_reached", 2: "dt_min_reached"}
MAXD = 64


class Problem(ctypes.Structure):
    _fields_ = [
        ("solver", ctypes.c_int32), ("controller", ctypes.c_int32), ("vf", ctypes.c_int32),
        ("d", ctypes.c_int32), ("hidden", ctypes.c_int32), ("n_ts", ctypes.c_int32),
        ("save_t0", ctypes.c_int32), ("save_t1", ctypes.c_int32),
        ("save_steps", ctypes.c_int32), ("max_steps", ctypes.c_int32),
        ("force_dtmin", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("t0", ctypes.c_double), ("dt0", ctypes.c_double),
        ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
        ("pcoeff", ctypes.c_double), ("icoeff", ctypes.c_double), ("dcoeff", ctypes.c_double),
        ("dtmin", ctypes.c_double), ("dtmax", ctypes.c_double),
        ("factormin", ctypes.c_double), ("factormax", ctypes.c_double),
        ("safety", ctypes.c_double), ("error_order", ctypes.c_double),
        ("step_ts", ctypes.c_void_p), ("jump_ts", ctypes.c_void_p),
        ("n_step_ts", ctypes.c_int32), ("n_jump_ts", ctypes.c_int32),
        ("event_coeffs", ctypes.c_void_p), ("event_flags", ctypes.c_void_p),
        ("event_mask", ctypes.c_void_p),
        ("n_events", ctypes.c_int32), ("event_root_find", ctypes.c_int32),
        ("adj_base_d", ctypes.c_int32), ("pad2_", ctypes.c_int32),
        ("grid_y", ctypes.c_void_p),
        ("rtol_b", ctypes.c_void_p), ("atol_b", ctypes.c_void_p),
    ]


_lib = None


def build():
    """Compile liberk_oracle.so with the committed Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        dp = P(ctypes.c_double)
        ip = P(ctypes.c_int32)
        L.orc_solve.argtypes = [P(Problem), ctypes.c_long, dp, dp, ctypes.c_long, dp,
                                ctypes.c_long, dp, ctypes.c_long, dp, dp, dp, ip, ip, dp,
                                ctypes.c_int]
        L.orc_solve.restype = ctypes.c_int
        L.orc_n_out.argtypes = [P(Problem)]
        L.orc_n_out.restype = ctypes.c_int
        L.orc_tableau.argtypes = [ctypes.c_int, dp, dp, dp, dp, ip]
        L.orc_single_step.argtypes = [P(Problem), dp, ctypes.c_double, ctypes.c_double,
                                      dp, dp, dp, dp]
        L.orc_param_count.argtypes = [P(Problem)]
        L.orc_param_count.restype = ctypes.c_long
        L.orc_backsolve.argtypes = [P(Problem), ctypes.c_long, dp, ctypes.c_long, dp, dp, dp,
                                    dp, dp, ip, ctypes.c_int]
        L.orc_backsolve.restype = ctypes.c_int
        L.orc_adjoint.argtypes = [P(Problem), ctypes.c_long, dp, ctypes.c_long, dp,
                                  ctypes.c_long, dp, ctypes.c_long, dp, dp, ip, dp, dp, dp,
                                  ctypes.c_int]
        L.orc_adjoint.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def tableau(solver):
    """(a[7,7], b[7], b_err[7], c[7], (S, fsal, ssal, order)) of the oracle's tableau."""
    a = np.zeros(49); b = np.zeros(7); be = np.zeros(7); c = np.zeros(7)
    info = np.zeros(4, dtype=np.int32)
    lib().orc_tableau(SOLVERS[solver], _dp(a), _dp(b), _dp(be), _dp(c), _ip(info))
    return a.reshape(7, 7), b, be, c, tuple(int(x) for x in info)


def make_problem(solver="tsit5", vf="lv", d=2, *, t0=0.0, dt0=None, rtol=1e-6, atol=1e-8,
                 controller="pid", n_ts=0, save_t0=False, save_t1=False, save_steps=False,
                 max_steps=4096, hidden=0, pcoeff=0.0, icoeff=1.0, dcoeff=0.0, dtmin=None,
                 dtmax=None, force_dtmin=True, factormin=0.2, factormax=10.0, safety=0.9,
                 error_order=None, step_ts=None, jump_ts=None, events=None,
                 event_root_find=False):
    """The problem struct.  rtol / atol may be [B] arrays (per-instance tolerances, as a
    jax.vmap over the PIDController's tolerances).  step_ts / jump_ts wrap the PID controller in
    ClipStepSizeController (sorted here; the arrays stay referenced by the struct).
    events: list of (c_t, c_y [d], c_0, boolean, direction) affine event conditions,
    direction None / True (up) / False (down)."""
    nan = float("nan")
    clip = {}
    for name, v in (("step_ts", step_ts), ("jump_ts", jump_ts)):
        arr = None if v is None else np.ascontiguousarray(np.sort(np.asarray(v, dtype=np.float64)))
        clip[name] = arr
    prob = Problem(
        solver=SOLVERS[solver], controller=CONTROLLERS[controller], vf=VFS[vf], d=d,
        hidden=hidden, n_ts=n_ts, save_t0=int(save_t0), save_t1=int(save_t1),
        save_steps=int(save_steps), max_steps=max_steps, force_dtmin=int(force_dtmin), pad_=0,
        t0=float(t0), dt0=nan if dt0 is None else float(dt0),
        rtol=float(rtol) if np.ndim(rtol) == 0 else 0.0,
        atol=float(atol) if np.ndim(atol) == 0 else 0.0,
        pcoeff=float(pcoeff), icoeff=float(icoeff), dcoeff=float(dcoeff),
        dtmin=nan if dtmin is None else float(dtmin), dtmax=nan if dtmax is None else float(dtmax),
        factormin=float(factormin), factormax=float(factormax), safety=float(safety),
        error_order=nan if error_order is None else float(error_order),
        step_ts=None if clip["step_ts"] is None else clip["step_ts"].ctypes.data,
        jump_ts=None if clip["jump_ts"] is None else clip["jump_ts"].ctypes.data,
        n_step_ts=0 if clip["step_ts"] is None else clip["step_ts"].size,
        n_jump_ts=0 if clip["jump_ts"] is None else clip["jump_ts"].size)
    for name, v in (("rtol_b", rtol), ("atol_b", atol)):   # per-instance tolerances [B]
        if np.ndim(v) > 0:
            arr = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
            setattr(prob, name, arr.ctypes.data)
            clip[name] = arr
    if events:
        co = np.ascontiguousarray([[float(e[0])] + [float(v) for v in e[1]] + [float(e[2])]
                                   for e in events], dtype=np.float64)
        fl = np.ascontiguousarray(
            [int(bool(e[3])) | ((0 if e[4] is None else (1 if e[4] else 2)) << 1) for e in events],
            dtype=np.int32)
        prob.event_coeffs = co.ctypes.data
        prob.event_flags = fl.ctypes.data
        prob.n_events = len(events)
        prob.event_root_find = int(bool(event_root_find))
        clip["events"] = (co, fl)
    prob._keep = clip
    return prob


def solve(y0, t1, params, *, ts=None, dW=None, record_grid=False, nthreads=0, **kw):
    """Batched FP64 solve.

    y0: [B, d]; t1: scalar or [B]; params: [P] shared or [B, P] per instance;
    ts: None, [n_ts] shared or [B, n_ts] per instance; dW: [steps, B, d] or None.
    Returns dict with ys [B, n_out, d], ts [B, n_out], stats [B, 3], result [B]
    and, if record_grid, grid [B, max_steps, 2] (normalised accepted step intervals).
    """
    y0 = np.ascontiguousarray(np.atleast_2d(np.asarray(y0, dtype=np.float64)))
    B, d = y0.shape
    t1 = np.asarray(t1, dtype=np.float64)
    t1_stride = 1 if t1.ndim == 1 else 0
    t1 = np.ascontiguousarray(np.atleast_1d(t1))
    params = np.asarray(params, dtype=np.float64)
    param_stride = params.shape[-1] if params.ndim == 2 else 0
    params = np.ascontiguousarray(params)
    if ts is not None:
        ts = np.asarray(ts, dtype=np.float64)
        n_ts = ts.shape[-1]
        ts_stride = n_ts if ts.ndim == 2 else 0
        ts = np.ascontiguousarray(ts)
    else:
        n_ts, ts_stride = 0, 0
    p = make_problem(d=d, n_ts=n_ts, **kw)
    L = lib()
    n_out = L.orc_n_out(ctypes.byref(p))
    ys = np.empty((B, n_out, d))
    tso = np.empty((B, n_out))
    stats = np.empty((B, 3), dtype=np.int32)
    res = np.empty(B, dtype=np.int32)
    # entries past each instance's accepted-step count are left unset (lazily allocated)
    grid = np.empty((B, p.max_steps, 2)) if record_grid else None
    grid_y = np.empty((B, p.max_steps, d)) if record_grid else None
    if record_grid:
        p.grid_y = grid_y.ctypes.data
    if dW is not None:
        dW = np.ascontiguousarray(np.asarray(dW, dtype=np.float64))
    ev_mask = None
    if p.n_events > 0:
        ev_mask = np.zeros((B, p.n_events), dtype=np.int32)
        p.event_mask = ev_mask.ctypes.data
    rc = L.orc_solve(ctypes.byref(p), B, _dp(y0), _dp(t1), t1_stride, _dp(ts), ts_stride,
                     _dp(params), param_stride, _dp(dW), _dp(ys), _dp(tso), _ip(stats),
                     _ip(res), _dp(grid), int(nthreads))
    if rc != 0:
        raise ValueError(f"oracle rejected the problem (code {rc})")
    out = dict(ys=ys, ts=tso, stats=stats, result=res)
    if ev_mask is not None:
        out["event_mask"] = ev_mask.astype(bool)
    if record_grid:
        out["grid"] = grid
        out["grid_y"] = grid_y
    return out


def adjoint(t1, params, fwd, ys_bar, *, ts=None, nthreads=0, **kw):
    """Gradients of sum(ys * ys_bar) by the discrete adjoint (reverse mode through the
    accepted steps, RecursiveCheckpointAdjoint) for the forward problem described by
    t1 / params / ts / **kw (as for solve), whose solve(..., record_grid=True) result is
    fwd.  Returns (y0_bar [B, d], params_bar [B, P])."""
    grid, grid_y, stats = fwd["grid"], fwd["grid_y"], fwd["stats"]
    B, _, d = grid_y.shape
    t1 = np.asarray(t1, dtype=np.float64)
    t1_stride = 1 if t1.ndim == 1 else 0
    t1 = np.ascontiguousarray(np.atleast_1d(t1))
    params = np.asarray(params, dtype=np.float64)
    param_stride = params.shape[-1] if params.ndim == 2 else 0
    params = np.ascontiguousarray(params)
    if ts is not None:
        ts = np.ascontiguousarray(np.asarray(ts, dtype=np.float64))
        n_ts, ts_stride = ts.shape[-1], (ts.shape[-1] if ts.ndim == 2 else 0)
    else:
        n_ts, ts_stride = 0, 0
    p = make_problem(d=d, n_ts=n_ts, **kw)
    L = lib()
    P = int(L.orc_param_count(ctypes.byref(p)))
    y0b = np.zeros((B, d))
    qb = np.zeros((B, P))
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    grid_y = np.ascontiguousarray(grid_y, dtype=np.float64)
    stats = np.ascontiguousarray(stats, dtype=np.int32)
    ys_bar = np.ascontiguousarray(ys_bar, dtype=np.float64)
    rc = L.orc_adjoint(ctypes.byref(p), B, _dp(params), param_stride, _dp(t1), t1_stride,
                       _dp(ts), ts_stride, _dp(grid), _dp(grid_y), _ip(stats), _dp(ys_bar),
                       _dp(y0b), _dp(qb), int(nthreads))
    if rc != 0:
        raise ValueError(f"oracle rejected the adjoint problem (code {rc})")
    return y0b, qb


def single_step(solver, vf, y0, ts0, ts1, params, hidden=0):
    y0 = np.ascontiguousarray(np.asarray(y0, dtype=np.float64))
    d = y0.shape[0]
    p = make_problem(solver=solver, vf=vf, d=d, hidden=hidden)
    params = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
    y1 = np.empty(d); ye = np.empty(d); k = np.empty((7, d))
    lib().orc_single_step(ctypes.byref(p), _dp(params), float(ts0), float(ts1), _dp(y0),
                          _dp(y1), _dp(ye), _dp(k))
    return y1, ye, k


def mlp_pack(W1, b1, W2, b2, W3, b3):
    """Flatten MLP weights in the oracle's parameter order."""
    return np.concatenate([np.asarray(x, dtype=np.float64).ravel()
                           for x in (W1, b1, W2, b2, W3, b3)])


def backsolve(y0, t1, params, ys, ts_out, ys_bar, *, ts=None, nthreads=0, **kw):
    """Gradients of sum(ys * ys_bar) by the continuous adjoint (BacksolveAdjoint) for the
    forward problem described by y0 / t1 / params / ts / **kw (as for solve), whose saved
    states and times are ys / ts_out.  Returns (y0_bar [B, d], params_bar [B, P], result)."""
    y0 = np.ascontiguousarray(np.atleast_2d(np.asarray(y0, dtype=np.float64)))
    B, d = y0.shape
    params = np.asarray(params, dtype=np.float64)
    param_stride = params.shape[-1] if params.ndim == 2 else 0
    params = np.ascontiguousarray(params)
    n_ts = 0 if ts is None else np.asarray(ts).shape[-1]
    p = make_problem(d=d, n_ts=n_ts, **kw)
    L = lib()
    P = int(L.orc_param_count(ctypes.byref(p)))
    y0b = np.zeros((B, d))
    qb = np.zeros((B, P))
    res = np.zeros(B, dtype=np.int32)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    ts_out = np.ascontiguousarray(ts_out, dtype=np.float64)
    ys_bar = np.ascontiguousarray(ys_bar, dtype=np.float64)
    rc = L.orc_backsolve(ctypes.byref(p), B, _dp(params), param_stride, _dp(ys), _dp(ts_out),
                         _dp(ys_bar), _dp(y0b), _dp(qb), _ip(res), int(nthreads))
    if rc != 0:
        raise ValueError(f"oracle rejected the backsolve problem (code {rc})")
    return y0b, qb, res


def max_rel_err(y, y_ref, atol):
    """max |y - y_ref| / max(|y_ref|, atol) over finite reference entries."""
    y = np.asarray(y, dtype=np.float64)
    y_ref = np.asarray(y_ref, dtype=np.float64)
    m = np.isfinite(y_ref)
    if not m.any():
        return 0.0
    return float(np.max(np.abs(y[m] - y_ref[m]) / np.maximum(np.abs(y_ref[m]), atol)))


__all__ = ["solve", "single_step", "tableau", "make_problem", "mlp_pack", "max_rel_err",
           "build", "lib", "SOLVERS", "VFS", "RESULT_NAMES"]
_ = math
