"""RESULTS and Solution (mirrors diffrax/_solution.py:13-132)."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional


class _Result(int):
    """An integer result code carrying its diffrax name and message."""

    def __new__(cls, value, name, message):
        obj = int.__new__(cls, value)
        obj.name = name
        obj.message = message
        obj._value = value
        return obj

    def __repr__(self):
        return f"RESULTS.{self.name}"


class _ResultsMeta(type):
    def __getitem__(cls, item):
        return cls._by_value[int(item)].message

    def __len__(cls):
        return len(cls._by_value)

    def __iter__(cls):
        return iter(cls._by_value.values())


class RESULTS(metaclass=_ResultsMeta):
    """Per-solve outcome codes.  Names and messages follow diffrax (_solution.py:13-31);
    successful == 0 as the reference tests pin (test_saveat_solution.py:21)."""

    successful = _Result(0, "successful", "")
    max_steps_reached = _Result(
        1, "max_steps_reached",
        "The maximum number of solver steps was reached. Try increasing `max_steps`.")
    dt_min_reached = _Result(
        2, "dt_min_reached",
        "The minimum step size was reached in the differential equation solver.")
    event_occurred = _Result(
        3, "event_occurred",
        "Terminating differential equation solve because an event occurred.")
    max_steps_rejected = _Result(
        4, "max_steps_rejected",
        "Maximum number of rejected steps was reached. Consider increasing "
        "`diffrax.ClipStepSizeController(store_rejected_steps==...)`.")
    internal_error = _Result(
        5, "internal_error",
        "An internal error occurred in Diffrax. This is a bug! Please open a GitHub "
        "issue with a minimum working example. (<50 lines of code is ideal)")
    _by_value: dict = {}


RESULTS._by_value = {int(r): r for r in (
    RESULTS.successful, RESULTS.max_steps_reached, RESULTS.dt_min_reached,
    RESULTS.event_occurred, RESULTS.max_steps_rejected, RESULTS.internal_error)}


def is_okay(result) -> bool:
    return int(result) in (0, 3)


def is_successful(result) -> bool:
    return int(result) == 0


@dataclass
class Solution:
    """Output of diffeqsolve (diffrax/_solution.py:82-132).

    Batched solves carry a leading batch axis on ts, ys, result and every stats entry
    except max_steps, exactly as jax.vmap over diffeqsolve would."""

    t0: Any
    t1: Any
    ts: Any
    ys: Any
    interpolation: Any = None
    stats: dict = field(default_factory=dict)
    result: Any = RESULTS.successful
    solver_state: Optional[Any] = None
    controller_state: Optional[Any] = None
    made_jump: Optional[Any] = None
    event_mask: Optional[Any] = None

    def evaluate(self, t0, t1=None, left: bool = True):
        if self.interpolation is None:
            raise ValueError(
                "Dense solution has not been saved; pass SaveAt(dense=True).")
        return self.interpolation.evaluate(t0, t1, left)

    def derivative(self, t, left: bool = True):
        if self.interpolation is None:
            raise ValueError(
                "Dense solution has not been saved; pass SaveAt(dense=True).")
        return self.interpolation.derivative(t, left)
