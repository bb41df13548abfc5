"""Pins of the oracle's continuous adjoint (BacksolveAdjoint, _adjoint.py:539-873 with
AdjointTerm, _term.py:767-917) on the CPU.

- test_against (test_adjoint.py:34-214): y' = A y with A = [[0.1, -1], [1, -0.1]] (the
  reference's _VectorField with its args), t0 = 0.3, t1 = 9.5, dt0 = None, Tsit5 with
  PIDController(rtol=atol=1e-8), loss = sum(ys) over the same SaveAt(t0, t1, ts) grid.
  The gradients are checked against the closed form -- dL/dy0 = sum_k expm(A^T dt_k) 1
  and dL/dA = sum_k of the Van Loan integral int_0^dt_k expm(A^T (dt_k - s)) 1 y0^T
  expm(A^T s) ds -- at the reference's tolerance for BacksolveAdjoint (atol 1e-5).
  A transposed or sign-flipped vector-Jacobian product fails it (A is not symmetric).
- Lotka-Volterra and Van der Pol (nonlinear vjps): against central differences of the
  oracle's forward solve at tight tolerances, per-instance parameters, forward and
  reverse time.
- the constant-step path (Heun, Dopri5) against the closed form.
- SaveAt(t0=True) alone: the gradient is the cotangent of the t0 row, parameters none.
"""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle

A_REF = np.array([[0.1, -1.0], [1.0, -0.1]])
Y0_REF = np.array([0.9, 5.4])


def _closed_form(A, y0, t0, times, w):
    """Exact gradients of sum_k w_k . y(times[k]) for y' = A y, y(t0) = y0."""
    d = A.shape[0]
    gy = np.zeros(d)
    gA = np.zeros((d, d))
    for tk, wk in zip(times, w):
        dt = tk - t0
        gy += expm(A.T * dt) @ wk
        M = np.zeros((2 * d, 2 * d))
        M[:d, :d] = A.T
        M[d:, d:] = A.T
        M[:d, d:] = np.outer(wk, y0)
        gA += expm(M * dt)[:d, d:]
    return gy, gA


def _fwd_and_back(y0, t1, params, ts, **kw):
    f = oracle.solve(y0, t1, params, ts=ts, **kw)
    assert (f["result"] == 0).all()
    w = np.ones_like(f["ys"])
    yb, qb, res = oracle.backsolve(y0, t1, params, f["ys"], f["ts"], w, ts=ts, **kw)
    assert (res == 0).all()
    return f, yb, qb


SAVEATS = [(t0, t1, ts) for t0 in (True, False) for t1 in (True, False)
           for ts in (None, [0.3], [2.0], [9.5], [1.0, 7.0], [0.3, 7.0, 9.5])
           if not (t0 is False and t1 is False and ts is None)]


@pytest.mark.parametrize("save_t0,save_t1,ts", SAVEATS)
def test_against_closed_form(save_t0, save_t1, ts):
    kw = dict(solver="tsit5", vf="linear", t0=0.3, dt0=None, rtol=1e-8, atol=1e-8,
              save_t0=save_t0, save_t1=save_t1)
    f, yb, qb = _fwd_and_back(Y0_REF[None], 9.5, A_REF.ravel(),
                              None if ts is None else np.array(ts), **kw)
    times = f["ts"][0]
    gy, gA = _closed_form(A_REF, Y0_REF, 0.3, times, np.ones((times.size, 2)))
    np.testing.assert_allclose(yb[0], gy, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(qb[0].reshape(2, 2), gA, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("solver,dt", [("heun", 1e-3), ("dopri5", 1e-2), ("tsit5", 1e-2)])
def test_constant_steps_closed_form(solver, dt):
    A = np.array([[-0.3, 0.8, 0.1], [-0.7, -0.2, 0.0], [0.2, 0.1, -0.5]])
    y0 = np.array([1.0, -0.5, 0.25])
    ts = np.array([0.5, 1.25, 2.0])
    f, yb, qb = _fwd_and_back(y0[None], 2.0, A.ravel(), ts, solver=solver, vf="linear",
                              dt0=dt, controller="constant")
    gy, gA = _closed_form(A, y0, 0.0, ts, np.ones((3, 3)))
    tol = 1e-6 if solver == "heun" else 1e-9
    np.testing.assert_allclose(yb[0], gy, rtol=tol, atol=tol)
    np.testing.assert_allclose(qb[0].reshape(3, 3), gA, rtol=tol, atol=tol)


def _fd(fun, x, h):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e.flat[i] = h
        g.flat[i] = (fun(x + e) - fun(x - e)) / (2 * h)
    return g


@pytest.mark.parametrize("vf,P", [("lv", 4), ("vdp", 1)])
@pytest.mark.parametrize("t0,t1", [(0.0, 4.0), (4.0, 0.0)])
def test_nonlinear_against_finite_differences(vf, P, t0, t1):
    rng = np.random.default_rng(7)
    B = 3
    y0 = rng.uniform(0.5, 2.0, size=(B, 2))
    q = (np.tile([1.5, 1.0, 3.0, 1.0], (B, 1)) * rng.uniform(0.8, 1.2, size=(B, 4))
         if vf == "lv" else rng.uniform(0.5, 2.0, size=(B, 1)))
    ts = np.linspace(t0, t1, 9)[1:]
    kw = dict(solver="dopri5", vf=vf, t0=t0, dt0=0.05 if t1 > t0 else -0.05, rtol=1e-11,
              atol=1e-11, save_t0=True, max_steps=100000)
    f = oracle.solve(y0, t1, q, ts=ts, **kw)
    w = rng.normal(size=f["ys"].shape)
    yb, qb, res = oracle.backsolve(y0, t1, q, f["ys"], f["ts"], w, ts=ts, **kw)
    assert (res == 0).all()
    for b in range(B):
        def L_y(y):
            return float((oracle.solve(y[None], t1, q[b], ts=ts, **kw)["ys"][0] * w[b]).sum())

        def L_q(p):
            return float((oracle.solve(y0[b][None], t1, p, ts=ts, **kw)["ys"][0] * w[b]).sum())

        # step-size decisions move with the perturbation: FD noise ~ tol / h = 1e-7
        for got, fd in ((yb[b], _fd(L_y, y0[b], 1e-4)), (qb[b], _fd(L_q, q[b], 1e-4))):
            assert np.max(np.abs(got - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd)))


def test_save_t0_only():
    y0 = np.array([[1.0, 2.0]])
    q = np.array([1.5, 1.0, 3.0, 1.0])
    f = oracle.solve(y0, 3.0, q, solver="tsit5", vf="lv", dt0=0.1, save_t0=True)
    w = np.array([[[0.5, -2.0]]])
    yb, qb, res = oracle.backsolve(y0, 3.0, q, f["ys"], f["ts"], w, solver="tsit5", vf="lv",
                                   dt0=0.1, save_t0=True)
    assert np.array_equal(yb, w[:, 0]) and np.array_equal(qb, np.zeros((1, 4)))
    assert res[0] == 0


def test_rejects_unsupported():
    y0 = np.ones((1, 2))
    with pytest.raises(ValueError):   # SaveAt(steps=True)
        oracle.backsolve(y0, 1.0, np.ones(4), np.ones((1, 4097, 2)), np.ones((1, 4097)),
                         np.ones((1, 4097, 2)), vf="lv", dt0=0.1, save_steps=True)


# ---- pyoracle.backsolve: the generic continuous adjoint that checks the neural-ODE kernel

from oracle import pyoracle  # noqa: E402


def _lv_aug(q):
    """Augmented Lotka-Volterra field (f, vjp_y(-a), vjp_q(-a)) over numpy float64."""
    def f(t, z):
        x, v = z[0], z[1]
        w0, w1 = -z[2], -z[3]
        fy = [q[0] * x - q[1] * x * v, q[3] * x * v - q[2] * v]
        gy = [w0 * (q[0] - q[1] * v) + w1 * (q[3] * v), w0 * (-(q[1] * x)) + w1 * (q[3] * x - q[2])]
        gq = [w0 * x, -(w0 * (x * v)), -(w1 * v), w1 * (x * v)]
        return np.array(fy + gy + gq)
    return f


def test_pyoracle_backsolve_matches_c_oracle():
    """pyoracle.backsolve (full norm, fused sums) against orc_backsolve on Lotka-Volterra:
    the same algorithm written twice agrees to rounding."""
    y0 = np.array([1.2, 0.7])
    q = np.array([1.5, 1.0, 3.0, 1.0])
    ts = np.array([0.5, 1.5, 3.0])
    kw = dict(solver="tsit5", vf="lv", t0=0.0, dt0=0.1, rtol=1e-7, atol=1e-9, save_t0=True,
              save_t1=True)
    f = oracle.solve(y0[None], 3.5, q, ts=ts, **kw)
    w = np.random.default_rng(3).normal(size=f["ys"].shape)
    yb, qb, res = oracle.backsolve(y0[None], 3.5, q, f["ys"], f["ts"], w, ts=ts, **kw)
    ay, aq, worst, _ = pyoracle.backsolve(_lv_aug(q), 2, 4, "tsit5", 0.0, -0.1, f["ts"][0],
                                          f["ys"][0], w[0], save_t0=True, rtol=1e-7, atol=1e-9)
    assert worst == 0 and res[0] == 0
    np.testing.assert_allclose(ay, yb[0], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(aq, qb[0], rtol=1e-11, atol=1e-12)


def _linear_aug(A):
    d = A.shape[0]

    def f(t, z):
        y, w = z[:d], -z[d:2 * d]
        return np.concatenate([A @ y, A.T @ w, np.outer(w, y).ravel()])
    return f


@pytest.mark.parametrize("seminorm", [False, True])
def test_pyoracle_backsolve_seminorm_closed_form(seminorm):
    """The seminorm leaves a_A out of the error control but not out of the solution: both
    norms give the closed-form gradient at the reference's backsolve tolerance, and the
    seminorm takes no more steps (its norm is over a subset of the components)."""
    ts = np.array([1.0, 4.0, 9.5])
    f = oracle.solve(Y0_REF[None], 9.5, A_REF.ravel(), ts=ts, solver="tsit5", vf="linear",
                     t0=0.3, dt0=0.1, rtol=1e-8, atol=1e-8)
    w = np.ones_like(f["ys"][0])
    runs = {}
    for sn in (False, True):
        runs[sn] = pyoracle.backsolve(_linear_aug(A_REF), 2, 4, "tsit5", 0.3, -0.1, f["ts"][0],
                                      f["ys"][0], w, save_t0=False, rtol=1e-8, atol=1e-8,
                                      seminorm=sn, fused=False)
    ay, aq, worst, steps = runs[seminorm]
    gy, gA = _closed_form(A_REF, Y0_REF, 0.3, f["ts"][0], w)
    assert worst == 0
    np.testing.assert_allclose(ay, gy, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(aq.reshape(2, 2), gA, rtol=1e-5, atol=1e-5)
    n_semi = sum(s[0] for s in runs[True][3])
    n_full = sum(s[0] for s in runs[False][3])
    assert n_semi <= n_full


def test_pyoracle_backsolve_reference_seminorm_recipe():
    """test_adjoint.py::test_adjoint_seminorm: y' = -y, y0 = 2 on [0, 1], Tsit5, forward
    PIDController(1e-3, 1e-6), backward the same tolerances with the seminorm; the gradient
    of y(1) is exp(-1)."""
    fwd = pyoracle.solve(lambda t, y: -y, "tsit5", 0.0, 1.0, None, np.array([2.0]),
                         rtol=1e-3, atol=1e-6, save_t1=True)
    aug = lambda t, z: np.array([-z[0], z[1]])    # (f, vjp_y(-a)); no parameters
    ay, aq, worst, _ = pyoracle.backsolve(aug, 1, 0, "tsit5", 0.0, None, fwd["ts"], fwd["ys"],
                                          np.ones((1, 1)), save_t0=False, rtol=1e-3,
                                          atol=1e-6, seminorm=True)
    assert worst == 0 and aq.size == 0
    assert abs(ay[0] - np.exp(-1.0)) < 1e-3
