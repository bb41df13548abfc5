"""Event (mirrors diffrax/_event.py:13-118) for the small-system kernel.

cond_fn must be a registered device condition: AffineCondition, c_t t + c_y . y + c_0,
real-valued (the solve stops when it changes sign, optionally only upwards or downwards)
or boolean (stops on the first step where it is > 0).  As in the reference, cond_fn sees
the direction-normalised time.  With a root finder the event time is located on the last
step's dense interpolant; the device uses bisection to adjacent floats (the reference runs
optimistix.Newton with the finder's tolerances on the same interpolant)."""
from __future__ import annotations

from typing import Optional, Sequence

import torch


class AffineCondition:
    """cond_fn(t, y, args) = t_coeff * t + sum_j y_coeffs[j] * y[j] + offset
    (boolean=True: that value > 0)."""

    def __init__(self, t: float = 0.0, y: Optional[Sequence[float]] = None,
                 offset: float = 0.0, boolean: bool = False):
        self.t_coeff = float(t)
        self.y_coeffs = None if y is None else [float(v) for v in y]
        self.offset = float(offset)
        self.boolean = bool(boolean)

    def coeffs(self, d):
        yc = self.y_coeffs if self.y_coeffs is not None else [0.0] * d
        if len(yc) != d:
            raise ValueError(f"condition has {len(yc)} state coefficients, the state has {d}")
        return [self.t_coeff] + list(yc) + [self.offset]

    def __call__(self, t, y, args=None, **kwargs):
        yt = torch.as_tensor(y)
        v = self.t_coeff * t + self.offset
        if self.y_coeffs is not None:
            v = v + (yt * torch.as_tensor(self.y_coeffs, dtype=yt.dtype)).sum(-1)
        return v > 0 if self.boolean else v


class Newton:
    """Stands for optimistix.Newton(rtol, atol, norm) as Event.root_finder: its presence
    asks for the event time to be located (the kernel's finder is bisection)."""

    def __init__(self, rtol: float = 1e-5, atol: float = 1e-5, norm=None):
        self.rtol, self.atol, self.norm = rtol, atol, norm


class Event:
    """Terminate the solve when a condition triggers (_event.py:13-56)."""

    def __init__(self, cond_fn, root_finder=None, direction=None):
        conds = list(cond_fn) if isinstance(cond_fn, (list, tuple)) else [cond_fn]
        if direction in (None, False, True):
            dirs = [direction] * len(conds)
        else:
            dirs = list(direction)
            if len(dirs) != len(conds):
                raise ValueError("Missmatch in the structure of `cond_fn` and `direction`.")
        if any(x not in (None, False, True) for x in dirs):
            raise ValueError(
                "`direction` must be a `None`, `bool`, or a PyTree of `None | bool`s "
                "with the same structure as `cond_fn`.")
        for c in conds:
            if not isinstance(c, AffineCondition):
                raise NotImplementedError(
                    "diffrax_amd evaluates registered event conditions on the device: "
                    "use AffineCondition")
        if len(conds) > 4:
            raise NotImplementedError("at most 4 event conditions")
        self.cond_fn = cond_fn
        self.conds = conds
        self.direction = dirs
        self.root_finder = root_finder
        self.is_list = isinstance(cond_fn, (list, tuple))

    def flags(self):
        return [int(c.boolean) | ((0 if dr is None else (1 if dr else 2)) << 1)
                for c, dr in zip(self.conds, self.direction)]
