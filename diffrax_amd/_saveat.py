"""SaveAt / SubSaveAt (mirrors diffrax/_saveat.py:10-105)."""
from __future__ import annotations

from typing import Any, Callable, Optional


def save_y(t, y, args):
    return y


class SubSaveAt:
    """Which times to save: t0, t1, ts and/or every accepted step (_saveat.py:14-48)."""

    def __init__(self, *, t0: bool = False, t1: bool = False, ts: Any = None,
                 steps: bool | int = False, fn: Callable = save_y):
        if not (t0 or t1 or (ts is not None) or steps):
            raise ValueError("Empty saveat -- nothing will be saved.")
        self.t0 = bool(t0)
        self.t1 = bool(t1)
        self.ts = ts
        self.steps = int(steps)
        self.fn = fn


class SaveAt:
    """Determines what to save from diffeqsolve (_saveat.py:65-105)."""

    def __init__(self, *, t0: bool = False, t1: bool = False, ts: Any = None,
                 steps: bool | int = False, fn: Callable = save_y,
                 subs: Optional[SubSaveAt] = None, dense: bool = False,
                 solver_state: bool = False, controller_state: bool = False,
                 made_jump: bool = False):
        if subs is None:
            if t0 or t1 or (ts is not None) or steps:
                subs = SubSaveAt(t0=t0, t1=t1, ts=ts, steps=steps, fn=fn)
        else:
            if t0 or t1 or (ts is not None) or steps:
                raise ValueError(
                    "Cannot pass both `subs` and any of `t0`, `t1`, `ts`, `steps` to "
                    "`SaveAt`.")
        self.subs = subs
        self.dense = bool(dense)
        self.solver_state = bool(solver_state)
        self.controller_state = bool(controller_state)
        self.made_jump = bool(made_jump)
