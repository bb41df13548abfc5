"""Pins of the oracle's discrete adjoint (orc_adjoint: reverse mode through the accepted
steps, the gradient of RecursiveCheckpointAdjoint, _adjoint.py:264-312) on the CPU.

- torch float64 autograd through oracle.pyoracle.replay of the same accepted-step grid
  (an independent restatement of the step and the save interpolation): every solver,
  Lotka-Volterra / Van der Pol / linear fields, SaveAt t0, ts, steps and t1, reverse
  time, per-instance save times, ClipStepSizeController grids, to 1e-12 relative;
- y' = A y against the closed form (expm, Van Loan block for dL/dA) at the solve's
  accuracy, so a transposed or sign-flipped step transpose fails;
- the zero-length solve: every row is y0.
"""
import numpy as np
import pytest
import torch
from scipy.linalg import expm

import oracle
from oracle import pyoracle


def _field(vf, qt):
    if vf == "lv":
        def f(t, z):
            x, v = z[0], z[1]
            return torch.stack([qt[0] * x - qt[1] * x * v, qt[3] * x * v - qt[2] * v])
    elif vf == "vdp":
        def f(t, z):
            x, v = z[0], z[1]
            return torch.stack([v, qt[0] * (1 - x * x) * v - x])
    else:
        d = int(round(np.sqrt(qt.numel())))

        def f(t, z):
            return qt.reshape(d, d) @ z
    return f


def _replay_grads(vf, solver, fwd, b, y0, q, ts, w, direction, save_t0, save_t1, save_steps):
    n = int(fwd["stats"][b, 1])
    grid = [tuple(g) for g in fwd["grid"][b, :n]]
    y0t = torch.tensor(y0, requires_grad=True)
    qt = torch.tensor(q, requires_grad=True)
    out = pyoracle.replay(_field(vf, qt), solver, grid, y0t, list(ts), direction=direction,
                          save_t1=save_t1, save_steps=save_steps)
    rows = ([y0t] if save_t0 else []) + out
    L = sum((r * torch.tensor(w[i])).sum() for i, r in enumerate(rows))
    gy, gq = torch.autograd.grad(L, (y0t, qt))
    return gy.numpy(), gq.numpy()


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


CASES = [("lv", 4), ("vdp", 1), ("linear", 9)]


@pytest.mark.parametrize("solver", ["tsit5", "dopri5", "heun"])
@pytest.mark.parametrize("vf,P", CASES)
@pytest.mark.parametrize("save", ["t0+ts+t1", "steps", "ts", "steps3+ts+t1"])
def test_against_autograd_replay(solver, vf, P, save):
    rng = np.random.default_rng([CASES.index((vf, P)), len(solver), len(save)])
    B = 3
    d = 3 if vf == "linear" else 2
    y0 = rng.uniform(0.5, 2.0, size=(B, d))
    if vf == "lv":
        q = np.array([1.5, 1.0, 3.0, 1.0]) * rng.uniform(0.8, 1.2, size=(B, 4))
    elif vf == "vdp":
        q = rng.uniform(0.2, 2.0, size=(B, 1))
    else:
        q = (rng.normal(size=(B, 3, 3)) * 0.5 - np.eye(3)).reshape(B, 9)
    ts = None if save == "steps" else np.sort(rng.uniform(0.0, 3.0, size=7))
    flags = dict(save_t0=save == "t0+ts+t1", save_t1=save in ("t0+ts+t1", "steps3+ts+t1"),
                 save_steps={"steps": 1, "steps3+ts+t1": 3}.get(save, 0))
    kw = dict(solver=solver, vf=vf, dt0=0.05, rtol=1e-3 if solver == "heun" else 1e-6,
              atol=1e-8, max_steps=2000, **flags)
    fwd = oracle.solve(y0, 3.0, q, ts=ts, record_grid=True, **kw)
    assert (fwd["result"] == 0).all()
    w = rng.normal(size=fwd["ys"].shape)
    w[~np.isfinite(fwd["ys"])] = 0.0           # SaveAt(steps) rows past the last step
    yb, qb = oracle.adjoint(3.0, q, fwd, w, ts=ts, **kw)
    for b in range(B):
        n = int(fwd["stats"][b, 1])
        k = flags["save_steps"]
        nrows = (flags["save_t0"] + (0 if ts is None else ts.size) + (n // k if k else 0)
                 + (flags["save_t1"] and not (k and n % k == 0)))
        t1_row = bool(flags["save_t1"] and not (k and n % k == 0))
        gy, gq = _replay_grads(vf, solver, fwd, b, y0[b], q[b], [] if ts is None else ts,
                               w[b, :nrows], 1, flags["save_t0"], t1_row, k)
        assert _rel(yb[b], gy) <= 1e-12
        assert _rel(qb[b], gq) <= 1e-12


def test_reverse_time_per_instance_ts_and_clip():
    rng = np.random.default_rng(9)
    B = 4
    y0 = rng.uniform(0.5, 2.0, size=(B, 2))
    q = np.array([1.5, 1.0, 3.0, 1.0])
    ts = np.sort(rng.uniform(0.0, 5.0, size=(B, 6)), axis=1)[:, ::-1].copy()
    kw = dict(solver="dopri5", vf="lv", t0=5.0, dt0=-0.1, rtol=1e-6, atol=1e-8, save_t1=True,
              step_ts=[1.0, 2.0, 3.0], jump_ts=[2.5])
    fwd = oracle.solve(y0, 0.0, q, ts=ts, record_grid=True, **kw)
    assert (fwd["result"] == 0).all()
    w = rng.normal(size=fwd["ys"].shape)
    yb, qb = oracle.adjoint(0.0, q, fwd, w, ts=ts, **kw)
    for b in range(B):
        gy, gq = _replay_grads("lv", "dopri5", fwd, b, y0[b], q, ts[b], w[b], -1, False,
                               True, False)
        assert _rel(yb[b], gy) <= 1e-12 and _rel(qb[b], gq) <= 1e-12


def test_linear_closed_form():
    A = np.array([[0.1, -1.0], [1.0, -0.1]])
    y0 = np.array([0.9, 5.4])
    ts = np.array([1.0, 7.0, 9.5])
    kw = dict(solver="tsit5", vf="linear", t0=0.3, rtol=1e-10, atol=1e-10, max_steps=20000)
    fwd = oracle.solve(y0[None], 9.5, A.ravel(), ts=ts, record_grid=True, **kw)
    w = np.ones_like(fwd["ys"])
    yb, qb = oracle.adjoint(9.5, A.ravel(), fwd, w, ts=ts, **kw)
    gy, gA = np.zeros(2), np.zeros((2, 2))
    for tk in ts:
        dt = tk - 0.3
        gy += expm(A.T * dt) @ np.ones(2)
        M = np.zeros((4, 4))
        M[:2, :2] = M[2:, 2:] = A.T
        M[:2, 2:] = np.outer(np.ones(2), y0)
        gA += expm(M * dt)[:2, 2:]
    np.testing.assert_allclose(yb[0], gy, rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(qb[0].reshape(2, 2), gA, rtol=1e-7, atol=1e-7)


def test_zero_length_solve():
    y0 = np.array([[1.0, 2.0]])
    q = np.array([1.5, 1.0, 3.0, 1.0])
    kw = dict(solver="tsit5", vf="lv", t0=1.0, dt0=0.1, save_t0=True, save_t1=True)
    ts = np.array([1.0, 1.0])
    fwd = oracle.solve(y0, 1.0, q, ts=ts, record_grid=True, **kw)
    w = np.arange(8.0).reshape(1, 4, 2)
    yb, qb = oracle.adjoint(1.0, q, fwd, w, ts=ts, **kw)
    assert np.array_equal(yb[0], w[0].sum(0)) and not qb.any()
