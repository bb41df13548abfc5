"""Pins of the oracle's Event restatement (_integrate.py:542-633, 686-821) with affine
conditions, on the CPU, following the reference's test_event.py recipes.  Vector fields
that the registered set lacks are written as linear systems: y' = 1 is (y, c)' = (c, 0)
with c = 1."""
import math

import numpy as np

import oracle

E = math.e
ONE = np.array([[0.0, 1.0], [0.0, 0.0]]).ravel()     # (y, c)' = (c, 0)


def _solve(y0, A, t1, events, root=True, **kw):
    base = dict(solver="tsit5", vf="linear", dt0=1.0, controller="constant", save_t1=True)
    base.update(kw)
    return oracle.solve(np.atleast_2d(y0), t1, np.asarray(A, dtype=float).ravel(),
                        events=events, event_root_find=root, **base)


def test_continuous_event_time():          # test_event.py:181-197
    r = _solve([1.0], [1.0], np.inf, [(0.0, [1.0], -E, False, None)])
    assert r["result"][0] == 3 and r["event_mask"][0, 0]
    assert np.isclose(r["ts"][0, 0], 1.0, rtol=1e-4)
    assert np.isclose(r["ys"][0, 0, 0], E, rtol=1e-12)   # the root of the interpolant


def test_continuous_event_value():         # test_event.py:200-216
    r = _solve([-10.0, 1.0], ONE, np.inf, [(0.0, [1.0, 0.0], 0.0, False, None)])
    assert abs(r["ys"][0, 0, 0]) <= 1e-5
    assert np.isclose(r["ts"][0, 0], 10.0, rtol=1e-12)


def test_continuous_no_event():            # test_event.py:219-236
    r = _solve([-10.0, 1.0], ONE, 5.0, [(0.0, [1.0, 0.0], 0.0, False, None)])
    assert r["result"][0] == 0 and not r["event_mask"][0, 0]
    assert r["ts"][0, 0] == 5.0


def test_continuous_two_events():          # test_event.py:239-262
    r = _solve([-10.0, 1.0], ONE, np.inf,
               [(1.0, [0.0, 0.0], -10.0, False, None), (0.0, [1.0, 0.0], 5.0, False, None)])
    assert list(r["event_mask"][0]) == [False, True]
    assert np.isclose(r["ts"][0, 0], 5.0, rtol=1e-5)


def test_terminate_without_root_finder():   # test_event.py:146-178
    r = _solve([1.0], [1.0], np.inf, [(1.0, [0.0], -10.0, True, None)], root=False)
    assert r["result"][0] == 3 and r["ys"][0, 0, 0] > 10.0 and r["ts"][0, 0] > 10.0
    r = _solve([1.0], [1.0], np.inf, [(1.0, [0.0], -10.0, False, None)], root=False)
    assert r["ts"][0, 0] >= 10.0


def test_event_before_the_first_step():    # test_event.py:124-143
    A = np.array([[0, 0, 1.0], [0, 0, 1.0], [0, 0, 0]])   # (a, x, c)' = (c, c, 0)
    r = _solve([1.0, -0.1, 1.0], A, 10.0, [(0.0, [0.0, -1.0, 0.0], 0.0, True, None)])
    assert r["stats"][0, 0] == 0 and r["result"][0] == 3
    assert np.array_equal(r["ys"][0, -1], [1.0, -0.1, 1.0]) and r["ts"][0, 0] == 0.0


def test_adaptive_stepping_event():        # test_event.py:319-351
    for y0 in (10.0, 100.0, 1000.0):
        r = _solve([y0], [-1.0], np.inf, [(0.0, [1.0], -1.0, False, None)],
                   controller="pid", rtol=1e-3, atol=1e-6)
        assert abs(r["ys"][0, 0, 0] - 1.0) <= 1e-5


def test_direction_and_saves_after_the_event_are_dropped():
    # Lotka-Volterra: x oscillates around 3; stop at the first upcrossing of x = 2.5
    q = np.array([1.5, 1.0, 3.0, 1.0])
    ts = np.linspace(0.0, 20.0, 41)
    kw = dict(solver="dopri5", vf="lv", dt0=0.1, rtol=1e-8, atol=1e-8, ts=ts, save_t1=True)
    up = oracle.solve(np.array([[1.0, 1.0]]), 20.0, q, events=[(0.0, [1.0, 0.0], -2.5, False, True)],
                      event_root_find=True, **kw)
    # rows: the ts saves up to the event, then the SaveAt(t1) row at the event, then inf
    # (the t1 row follows the surviving ts rows, as after the reference's unsave)
    n_fin = int(np.isfinite(up["ts"][0]).sum())
    te = up["ts"][0, n_fin - 1]
    assert np.isclose(up["ys"][0, n_fin - 1, 0], 2.5, atol=1e-12)
    assert np.array_equal(up["ts"][0, :n_fin - 1], ts[ts <= te])
    assert np.all(np.isinf(up["ts"][0, n_fin:])) and np.all(np.isinf(up["ys"][0, n_fin:]))
    # x rises through 2.5 there: the interpolant's x just before te is below 2.5
    full = oracle.solve(np.array([[1.0, 1.0]]), 20.0, q, **kw)
    xs = full["ys"][0, :-1, 0]
    k = int((ts < te).sum()) - 1
    assert xs[k] < 2.5 <= xs[k + 1]
    down = oracle.solve(np.array([[1.0, 1.0]]), 20.0, q,
                        events=[(0.0, [1.0, 0.0], -2.5, False, False)], event_root_find=True, **kw)
    n_dn = int(np.isfinite(down["ts"][0]).sum())
    assert down["ts"][0, n_dn - 1] > te    # the first downcrossing comes later


def test_steps_rows_with_root_finder():
    """SaveAt(steps=1) + root finder: rows are the accepted step ends before the event,
    then (t_event, y_event) (_integrate.py:797-822 unsave, 847-877 _save_t1 with
    pred = t1 | saved_via_steps)."""
    r = _solve([2.0], [-1.0], np.inf, [(0.0, [1.0], -1.0, False, None)], save_t1=False,
               save_steps=1, controller="pid", rtol=1e-3, atol=1e-6, dt0=0.1)
    ts = r["ts"][0][np.isfinite(r["ts"][0])]
    ys = r["ys"][0, :ts.size, 0]
    n = r["stats"][0, 1]
    assert ts.size == n                          # n - 1 step rows + the event row
    assert np.all(np.diff(ts) > 0) and np.isclose(ts[-1], np.log(2.0), rtol=5e-3)
    assert abs(ys[-1] - 1.0) <= 1e-5 and np.all(ys[:-1] > 1.0)
    plain = _solve([2.0], [-1.0], np.inf, [(0.0, [1.0], -1.0, False, None)], root=False,
                   save_t1=False, save_steps=1, controller="pid", rtol=1e-3, atol=1e-6, dt0=0.1)
    assert np.isfinite(plain["ts"][0]).sum() == n  # no root finder: the n step rows only
