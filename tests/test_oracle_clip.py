"""Pins of the oracle's ClipStepSizeController restatement (clip.py:120-428), on the CPU.

- _find_idx_with_hint: the first index strictly greater than t, whatever the hint
  (test_adaptive_stepsize_controller.py:307-317).
- _bump_next_t0: a step that ended on a listed time restarts one ULP after it, one that
  ended prevbefore it restarts one ULP after it too (two nextafters), anything else is
  unchanged (test_nested_clip_wrappers, ibid. 256-305).
- step_ts are stepped to exactly, forwards and backwards (test_step_ts, ibid. 18-43).
- jump_ts: the step before the jump ends at prevbefore(jump), the next starts at
  nextafter(jump); with step_ts alongside, both sets are hit and the solve succeeds
  (test_jump_ts, ibid. 46-89); a jump at t1 still saves at t1 (ibid. 343-365)."""
import ctypes

import numpy as np
import pytest

import oracle


def _find(ts, t, hint, direction=1):
    L = oracle.lib()
    L.orc_clip_find.restype = ctypes.c_int
    L.orc_clip_find.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_int]
    a = np.ascontiguousarray(ts, dtype=np.float64)
    return L.orc_clip_find(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.size,
                           direction, t, hint)


def _bump(ts, t, direction=1):
    L = oracle.lib()
    L.orc_clip_bump.restype = ctypes.c_double
    L.orc_clip_bump.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    a = np.ascontiguousarray(ts, dtype=np.float64)
    mj = ctypes.c_int(0)
    r = L.orc_clip_bump(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.size, direction, t,
                        ctypes.byref(mj))
    return r, bool(mj.value)


def test_find_idx_with_hint():
    ts = np.arange(5.0)
    for hint in (0, 2, 3, 5):
        assert _find(ts, 2.5, hint) == 3
        assert _find(ts, 2.0, hint) == 3   # the first value strictly greater
        assert _find(ts, 1.9, hint) == 2
    # wrapped for backward time: the view is -ts reversed = [-4, ..., 0]
    assert _find(ts, -2.5, 0, direction=-1) == 2


def test_bump_next_t0():
    ts = np.array([7.5, 23.0])
    assert _bump(ts, 23.0) == (np.nextafter(23.0, np.inf), True)
    assert _bump(ts, np.nextafter(7.5, -np.inf)) == (np.nextafter(7.5, np.inf), True)
    assert _bump(ts, 7.0) == (7.0, False)
    assert _bump(ts, -23.0, direction=-1) == (np.nextafter(-23.0, np.inf), True)


@pytest.mark.parametrize("backwards", [False, True])
def test_step_ts_are_stepped_to(backwards):
    t0, t1 = (5.0, 0.0) if backwards else (0.0, 5.0)
    r = oracle.solve(np.array([[1.0]]), t1, np.array([-0.2]), solver="dopri5", vf="linear",
                     t0=t0, rtol=1e-4, atol=1e-6, save_steps=True, step_ts=[3.0, 4.0])
    n = r["stats"][0, 1]
    ts = r["ts"][0, :n]
    assert r["result"][0] == 0
    assert 3.0 in ts and 4.0 in ts


@pytest.mark.parametrize("backwards", [False, True])
def test_jump_ts_stop_before_and_restart_after(backwards):
    t0, t1 = (15.0, 0.0) if backwards else (0.0, 15.0)
    dirn = -1.0 if backwards else 1.0
    kw = dict(solver="dopri5", vf="lv", t0=t0, rtol=1e-4, atol=1e-6, save_steps=True,
              record_grid=True)
    y0 = np.array([[1.0, 0.5]])
    q = np.array([1.5, 1.0, 3.0, 1.0])
    r = oracle.solve(y0, t1, q, jump_ts=[7.5], **kw)
    n = r["stats"][0, 1]
    g = r["grid"][0, :n]                      # accepted steps, direction-normalised time
    jn = 7.5 * dirn
    ends = g[:, 1]
    i = int(np.argmin(np.abs(ends - jn)))
    assert ends[i] == np.nextafter(jn, -np.inf)
    assert g[i + 1, 0] == np.nextafter(jn, np.inf)
    assert r["result"][0] == 0
    r = oracle.solve(y0, t1, q, jump_ts=[7.5], step_ts=[7.5], **kw)
    assert r["result"][0] == 0
    r = oracle.solve(y0, t1, q, jump_ts=[7.5], step_ts=[3.5, 8.0], **kw)
    ts = r["ts"][0, :r["stats"][0, 1]]
    assert r["result"][0] == 0 and 3.5 in ts and 8.0 in ts


def test_jump_at_t1_still_saves_t1():
    r = oracle.solve(np.array([[1.0]]), 1e3, np.array([-1.0]), solver="heun", vf="linear",
                     dt0=0.01, rtol=1e-6, atol=1e-6, ts=np.array([1e3]), jump_ts=[1e3])
    assert r["result"][0] == 0 and r["ts"][0, 0] == 1e3
