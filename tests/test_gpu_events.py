"""GPU tests of Event on the small-system kernel: the float64 form must equal the FP64
oracle (whose event logic tests/test_oracle_events.py pins against the reference's
test_event.py recipes) bit for bit -- saved rows, times, step statistics, result codes and
event masks -- for batches with per-instance initial states."""
import math

import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle

pytestmark = pytest.mark.gpu

ONE = np.array([[0.0, 1.0], [0.0, 0.0]])    # (y, c)' = (c, 0): y' = 1 with c = 1


def _both(y0, A_or_q, vf, t1, conds, dirs, root, saveat_kw, oracle_kw, ctrl, dt0,
          solver="tsit5"):
    y0 = np.asarray(y0, dtype=np.float64)
    cond = [dfx.AffineCondition(t=c[0], y=c[1], offset=c[2], boolean=c[3]) for c in conds]
    event = dfx.Event(cond if len(cond) > 1 else cond[0], dfx.Newton(1e-5, 1e-5) if root else None,
                      direction=dirs if len(cond) > 1 else dirs[0])
    if vf == "linear":
        term, args = dfx.ODETerm(dfx.LinearField()), torch.tensor(A_or_q, dtype=torch.float64)
        params = np.asarray(A_or_q, dtype=np.float64).ravel()
    else:
        term, args = dfx.ODETerm(dfx.LotkaVolterra()), tuple(A_or_q)
        params = np.asarray(A_or_q, dtype=np.float64)
    sol = dfx.diffeqsolve(term, getattr(dfx, solver.capitalize() if solver != "dopri5" else "Dopri5")(),
                          0.0, t1, dt0, torch.tensor(y0), args, saveat=dfx.SaveAt(**saveat_kw),
                          stepsize_controller=ctrl, event=event, max_steps=4096)
    ref = oracle.solve(y0, t1, params, solver=solver, vf=vf, dt0=dt0,
                       events=[(c[0], c[1], c[2], c[3], dr) for c, dr in zip(conds, dirs)],
                       event_root_find=root, **oracle_kw)
    return sol, ref


def _assert_equal(sol, ref):
    assert np.array_equal(sol.ys.cpu().numpy(), ref["ys"])
    assert np.array_equal(sol.ts.cpu().numpy(), ref["ts"])
    assert np.array_equal(sol.stats["num_steps"].cpu().numpy(), ref["stats"][:, 0])
    assert np.array_equal(sol.result.cpu().numpy(), ref["result"])


@pytest.mark.parametrize("constant", [True, False])
def test_event_batched_decay(cuda, constant):
    """test_event_vmap_y0 (test_event.py:354-392): y' = -y from y0 = 2..11, stop at y = 1."""
    y0 = (np.arange(10.0) + 2.0)[:, None]
    ctrl = dfx.ConstantStepSize() if constant else dfx.PIDController(rtol=1e-3, atol=1e-6)
    okw = dict(controller="constant") if constant else dict(rtol=1e-3, atol=1e-6)
    sol, ref = _both(y0, [[-1.0]], "linear", math.inf, [(0.0, [1.0], -1.0, False)], [None], True,
                     dict(t1=True), dict(save_t1=True, **okw), ctrl, 1.0)
    _assert_equal(sol, ref)
    assert np.all(np.abs(sol.ys.cpu().numpy()[:, -1, 0] - 1.0) <= 1e-5)
    assert (sol.result.cpu().numpy() == int(dfx.RESULTS.event_occurred)).all()
    assert sol.event_mask.cpu().numpy().all()


def test_event_two_conditions_and_no_event(cuda):
    """test_continuous_two_events / no_event (test_event.py:219-262), batched t1."""
    y0 = np.array([[-10.0, 1.0], [-10.0, 1.0], [-3.0, 1.0]])
    conds = [(1.0, [0.0, 0.0], -10.0, False), (0.0, [1.0, 0.0], 5.0, False)]
    t1 = torch.tensor([math.inf, 4.0, math.inf])
    y0t = torch.tensor(y0)
    ev = dfx.Event([dfx.AffineCondition(t=c[0], y=c[1], offset=c[2]) for c in conds], dfx.Newton())
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.LinearField()), dfx.Tsit5(), 0.0, t1, 1.0, y0t,
                          torch.tensor(ONE), saveat=dfx.SaveAt(t1=True), event=ev)
    ref = oracle.solve(y0, t1.numpy(), ONE.ravel(), solver="tsit5", vf="linear", dt0=1.0,
                       controller="constant", save_t1=True, event_root_find=True,
                       events=[(c[0], c[1], c[2], c[3], None) for c in conds])
    _assert_equal(sol, ref)
    mask = np.stack([m.cpu().numpy() for m in sol.event_mask], axis=-1)
    assert np.array_equal(mask, ref["event_mask"])
    assert list(mask[0]) == [False, True] and not mask[1].any()        # t1 = 4: no event
    assert list(mask[2]) == [True, False]                                # t = 10 comes first
    assert np.isclose(sol.ts.cpu().numpy()[0, 0], 5.0, rtol=1e-5)


@pytest.mark.parametrize("direction", [True, False, None])
def test_event_direction_and_dropped_saves(cuda, direction):
    y0 = np.random.default_rng(8).uniform(0.8, 1.6, size=(32, 2))
    ts = np.linspace(0.0, 20.0, 41)
    sol, ref = _both(y0, [1.5, 1.0, 3.0, 1.0], "lv", 20.0, [(0.0, [1.0, 0.0], -2.5, False)],
                     [direction], True, dict(ts=ts, t1=True),
                     dict(ts=ts, save_t1=True, rtol=1e-8, atol=1e-8), 
                     dfx.PIDController(rtol=1e-8, atol=1e-8), 0.1, solver="dopri5")
    _assert_equal(sol, ref)
    assert np.array_equal(sol.event_mask.cpu().numpy(), ref["event_mask"][:, 0])


def test_event_before_first_step_and_boolean_without_root_finder(cuda):
    A = np.array([[0, 0, 1.0], [0, 0, 1.0], [0, 0, 0]])    # (a, x, c)' = (c, c, 0)
    y0 = np.array([[1.0, -0.1, 1.0], [1.0, 0.5, 1.0]])
    sol, ref = _both(y0, A, "linear", 10.0, [(0.0, [0.0, -1.0, 0.0], 0.0, True)], [None], False,
                     dict(t1=True), dict(save_t1=True, controller="constant"),
                     dfx.ConstantStepSize(), 1.0)
    _assert_equal(sol, ref)
    st = sol.stats["num_steps"].cpu().numpy()
    assert st[0] == 0 and st[1] == 10          # instance 0 stops before stepping


@pytest.mark.parametrize("k,root", [(1, True), (1, False), (3, True)])
def test_event_with_saveat_steps(cuda, k, root):
    """SaveAt(steps=k) with an event: the step row after the event time is dropped and,
    with a root finder, the final (t_event, y_event) row is written even when the last
    step was saved via steps (_integrate.py:797-822, 847-877)."""
    y0 = (np.arange(8.0) + 2.0)[:, None]
    sol, ref = _both(y0, [[-1.0]], "linear", math.inf, [(0.0, [1.0], -1.0, False)], [None], root,
                     dict(steps=k), dict(save_steps=k, rtol=1e-3, atol=1e-6),
                     dfx.PIDController(rtol=1e-3, atol=1e-6), 0.1)
    _assert_equal(sol, ref)
    if root and k == 1:   # the event state closes the rows (k > 1: only if n % k == 0)
        ys = sol.ys.cpu().numpy()[..., 0]
        last = np.array([row[np.isfinite(row)][-1] for row in ys])
        assert np.all(np.abs(last - 1.0) <= 1e-5)
