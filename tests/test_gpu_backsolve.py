"""GPU tests of BacksolveAdjoint (_adjoint.py:539-873) on the small-system kernel, through
diffeqsolve + torch.autograd and the C ABI dfx_erk_backsolve.

- float64: y0 and parameter gradients bit-identical to the oracle's continuous adjoint
  (oracle.backsolve) on the same forward saves, for every registered field the adjoint
  covers and each solver, per-instance parameters, forward and reverse time, with and
  without SaveAt(t0);
- shared parameters: the per-instance adjoints summed over the batch (1e-13 relative:
  torch's reduction order);
- the reference's test_against recipe (test_adjoint.py:34-214): y' = A y, dt0 = None,
  PIDController(rtol=atol=1e-8), loss = sum(ys) over its SaveAt grid, against the closed
  form at the reference's tolerance (1e-5);
- BacksolveAdjoint(dt0=, solver=, stepsize_controller=, max_steps=) options reach the
  backward solves (bitwise against the oracle run with the same backward problem);
- float32-state form: within 2e-4 relative of the oracle's float64 adjoint."""
import numpy as np
import pytest
import torch
from scipy.linalg import expm

import diffrax_amd as dfx
import oracle

pytestmark = pytest.mark.gpu

SOLVERS = {"tsit5": dfx.Tsit5, "dopri5": dfx.Dopri5, "heun": dfx.Heun}
FIELDS = {"lv": (dfx.LotkaVolterra, 2, 4), "vdp": (dfx.VanDerPol, 2, 1),
          "linear2": (dfx.LinearField, 2, 4), "linear3": (dfx.LinearField, 3, 9)}


def _case(field, B, rng):
    cls, d, P = FIELDS[field]
    y0 = rng.uniform(0.5, 2.0, size=(B, d))
    if field == "lv":
        q = np.array([1.5, 1.0, 3.0, 1.0]) * rng.uniform(0.8, 1.2, size=(B, 4))
    elif field == "vdp":
        q = rng.uniform(0.2, 2.0, size=(B, 1))
    else:
        q = (rng.normal(size=(B, d, d)) * 0.5 - np.eye(d)).reshape(B, d * d)
    return cls, d, y0, q


def _args(field, qt, d):
    if field == "lv":
        return tuple(qt[:, i] for i in range(4))
    if field == "vdp":
        return qt[:, 0]
    return qt.reshape(-1, d, d)


def _grads(field, solver, y0, q, ts, t0, t1, dt0, w, save_t0, dtype=torch.float64,
           adjoint=None, controller=None, cuda=None):
    cls, d, _ = FIELDS[field]
    y0t = torch.tensor(y0, dtype=dtype, device=cuda, requires_grad=True)
    qt = torch.tensor(q, dtype=dtype, device=cuda, requires_grad=True)
    sol = dfx.diffeqsolve(dfx.ODETerm(cls()), SOLVERS[solver](), t0, t1, dt0, y0t,
                          _args(field, qt, d),
                          saveat=dfx.SaveAt(t0=save_t0, ts=ts),
                          stepsize_controller=controller or dfx.PIDController(rtol=1e-6,
                                                                              atol=1e-8),
                          adjoint=adjoint or dfx.BacksolveAdjoint(), max_steps=4096)
    wt = torch.tensor(w, dtype=dtype, device=cuda)
    gy, gq = torch.autograd.grad((sol.ys * wt).sum(), (y0t, qt))
    return sol, gy.cpu().numpy(), gq.cpu().numpy()


@pytest.mark.parametrize("solver", sorted(SOLVERS))
@pytest.mark.parametrize("field", sorted(FIELDS))
def test_bitwise_against_oracle(cuda, field, solver):
    rng = np.random.default_rng([sorted(FIELDS).index(field), sorted(SOLVERS).index(solver)])
    B = 96
    cls, d, y0, q = _case(field, B, rng)
    ts = np.array([0.5, 1.0, 2.25, 3.0])
    w = rng.normal(size=(B, 5, d))
    rtol = 1e-4 if solver == "heun" else 1e-6
    ctrl = dfx.PIDController(rtol=rtol, atol=1e-8)
    sol, gy, gq = _grads(field, solver, y0, q, ts, 0.0, 3.0, 0.05, w, True, controller=ctrl,
                         cuda=cuda)
    kw = dict(solver=solver, vf="linear" if field.startswith("linear") else field,
              dt0=0.05, rtol=rtol, atol=1e-8, save_t0=True)
    f = oracle.solve(y0, 3.0, q, ts=ts, **kw)
    assert np.array_equal(sol.ys.detach().cpu().numpy(), f["ys"])
    yb, qb, res = oracle.backsolve(y0, 3.0, q, f["ys"], f["ts"], w, ts=ts, **kw)
    assert (res == 0).all()
    assert np.array_equal(gy, yb)
    assert np.array_equal(gq, qb)


@pytest.mark.parametrize("save_t0", [False, True])
def test_reverse_time_and_uneven_t1(cuda, save_t0):
    rng = np.random.default_rng(31)
    B = 64
    _, d, y0, q = _case("lv", B, rng)
    t1 = rng.uniform(0.0, 2.0, size=B)          # every instance runs 5 -> t1[b] < 5
    y0t = torch.tensor(y0, dtype=torch.float64, device=cuda, requires_grad=True)
    qt = torch.tensor(q, dtype=torch.float64, device=cuda, requires_grad=True)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.LotkaVolterra()), dfx.Dopri5(), 5.0,
                          torch.tensor(t1), -0.1, y0t, _args("lv", qt, d),
                          saveat=dfx.SaveAt(t0=save_t0, t1=True),
                          stepsize_controller=dfx.PIDController(rtol=1e-7, atol=1e-9),
                          adjoint=dfx.BacksolveAdjoint())
    w = rng.normal(size=tuple(sol.ys.shape))
    gy, gq = torch.autograd.grad((sol.ys * torch.tensor(w, device=cuda)).sum(), (y0t, qt))
    kw = dict(solver="dopri5", vf="lv", t0=5.0, dt0=-0.1, rtol=1e-7, atol=1e-9,
              save_t0=save_t0, save_t1=True)
    f = oracle.solve(y0, t1, q, **kw)
    yb, qb, res = oracle.backsolve(y0, t1, q, f["ys"], f["ts"], w, **kw)
    assert (res == 0).all()
    assert np.array_equal(gy.cpu().numpy(), yb) and np.array_equal(gq.cpu().numpy(), qb)


def test_shared_parameters_sum_over_the_batch(cuda):
    rng = np.random.default_rng(2)
    B = 200
    y0 = rng.uniform(0.5, 2.0, size=(B, 2))
    q = np.array([1.5, 1.0, 3.0, 1.0])
    y0t = torch.tensor(y0, device=cuda, requires_grad=True)
    qt = torch.tensor(q, device=cuda, requires_grad=True)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.LotkaVolterra()), dfx.Tsit5(), 0.0, 4.0, 0.1, y0t,
                          tuple(qt[i] for i in range(4)),
                          saveat=dfx.SaveAt(ts=[1.0, 2.0, 4.0]),
                          stepsize_controller=dfx.PIDController(rtol=1e-6, atol=1e-8),
                          adjoint=dfx.BacksolveAdjoint())
    gq, = torch.autograd.grad(sol.ys.sum(), (qt,))
    kw = dict(solver="tsit5", vf="lv", dt0=0.1, rtol=1e-6, atol=1e-8)
    ts = np.array([1.0, 2.0, 4.0])
    f = oracle.solve(y0, 4.0, q, ts=ts, **kw)
    _, qb, _ = oracle.backsolve(y0, 4.0, q, f["ys"], f["ts"], np.ones_like(f["ys"]), ts=ts, **kw)
    np.testing.assert_allclose(gq.cpu().numpy(), qb.sum(0), rtol=1e-13)


def test_against_closed_form_reference_recipe(cuda):
    A = np.array([[0.1, -1.0], [1.0, -0.1]])
    y0 = np.array([0.9, 5.4])
    for st0 in (True, False):
        for st1 in (True, False):
            for ts in (None, [0.3], [2.0], [9.5], [1.0, 7.0], [0.3, 7.0, 9.5]):
                if not st0 and not st1 and ts is None:
                    continue
                y0t = torch.tensor(y0, device=cuda, requires_grad=True)
                At = torch.tensor(A, device=cuda, requires_grad=True)
                sol = dfx.diffeqsolve(dfx.ODETerm(dfx.LinearField()), dfx.Tsit5(), 0.3, 9.5,
                                      None, y0t, At, saveat=dfx.SaveAt(t0=st0, t1=st1, ts=ts),
                                      stepsize_controller=dfx.PIDController(rtol=1e-8,
                                                                            atol=1e-8),
                                      adjoint=dfx.BacksolveAdjoint())
                gy, gA = torch.autograd.grad(sol.ys.sum(), (y0t, At))
                ey, eA = np.zeros(2), np.zeros((2, 2))
                for tk in sol.ts.cpu().numpy():
                    dt = tk - 0.3
                    ey += expm(A.T * dt) @ np.ones(2)
                    M = np.zeros((4, 4))
                    M[:2, :2] = M[2:, 2:] = A.T
                    M[:2, 2:] = np.outer(np.ones(2), y0)
                    eA += expm(M * dt)[:2, 2:]
                np.testing.assert_allclose(gy.cpu().numpy(), ey, rtol=1e-5, atol=1e-5)
                np.testing.assert_allclose(gA.cpu().numpy(), eA, rtol=1e-5, atol=1e-5)


def test_backward_options(cuda):
    """BacksolveAdjoint(dt0=, solver=, stepsize_controller=, max_steps=) configure the
    backward solves only (_adjoint.py:623-632)."""
    rng = np.random.default_rng(4)
    B = 32
    _, d, y0, q = _case("vdp", B, rng)
    ts = np.array([1.0, 2.0])
    w = rng.normal(size=(B, 2, d))
    adj = dfx.BacksolveAdjoint(dt0=-0.02, solver=dfx.Dopri5(), max_steps=2000,
                               stepsize_controller=dfx.PIDController(rtol=1e-9, atol=1e-9))
    sol, gy, gq = _grads("vdp", "tsit5", y0, q, ts, 0.0, 2.0, 0.1, w, False, adjoint=adj,
                         cuda=cuda)
    f = oracle.solve(y0, 2.0, q, ts=ts, solver="tsit5", vf="vdp", dt0=0.1, rtol=1e-6,
                     atol=1e-8)
    # the oracle's backsolve takes the forward problem and negates its dt0
    yb, qb, res = oracle.backsolve(y0, 2.0, q, f["ys"], f["ts"], w, ts=ts, solver="dopri5",
                                   vf="vdp", dt0=0.02, rtol=1e-9, atol=1e-9, max_steps=2000)
    assert (res == 0).all()
    assert np.array_equal(gy, yb) and np.array_equal(gq, qb)


def test_float32_state_form(cuda):
    rng = np.random.default_rng(8)
    B = 128
    _, d, y0, q = _case("lv", B, rng)
    y0 = y0.astype(np.float32).astype(np.float64)
    q = q.astype(np.float32).astype(np.float64)
    ts = np.array([1.0, 2.0, 3.0])
    w = rng.normal(size=(B, 3, d)).astype(np.float32).astype(np.float64)
    _, gy, gq = _grads("lv", "tsit5", y0, q, ts, 0.0, 3.0, 0.1, w, False,
                       dtype=torch.float32, cuda=cuda)
    kw = dict(solver="tsit5", vf="lv", dt0=0.1, rtol=1e-6, atol=1e-8)
    f = oracle.solve(y0, 3.0, q, ts=ts, **kw)
    yb, qb, _ = oracle.backsolve(y0, 3.0, q, f["ys"], f["ts"], w, ts=ts, **kw)
    assert oracle.max_rel_err(gy, yb, 1.0) <= 2e-4
    assert oracle.max_rel_err(gq, qb, 1.0) <= 2e-4


def test_without_grad_is_a_plain_solve(cuda):
    """No input requires a gradient: BacksolveAdjoint changes nothing in the forward."""
    y0 = torch.rand(16, 2, dtype=torch.float64) + 0.5
    kw = dict(saveat=dfx.SaveAt(ts=[0.5, 1.0]),
              stepsize_controller=dfx.PIDController(rtol=1e-6, atol=1e-8))
    a = dfx.diffeqsolve(dfx.ODETerm(dfx.LotkaVolterra()), dfx.Tsit5(), 0.0, 1.0, 0.1, y0,
                        adjoint=dfx.BacksolveAdjoint(), **kw)
    b = dfx.diffeqsolve(dfx.ODETerm(dfx.LotkaVolterra()), dfx.Tsit5(), 0.0, 1.0, 0.1, y0, **kw)
    assert torch.equal(a.ys, b.ys)
