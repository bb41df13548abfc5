"""Step-size controllers (mirrors diffrax/_step_size_controller/{base,pid,constant}.py).

These are configuration objects: the accept/reject logic itself runs inside the device
kernels (erk_common.hpp), one lane per trajectory."""
from __future__ import annotations

import math
from typing import Callable, Optional

import numpy as np


def _tolerance(v, name):
    """A scalar tolerance, or a float64 [B] array of per-instance tolerances (what a
    jax.vmap over the controller's rtol / atol gives the reference)."""
    if hasattr(v, "detach"):
        v = v.detach().cpu().numpy()
    if np.ndim(v) == 0:
        return float(v)
    a = np.asarray(v, dtype=np.float64)
    if a.ndim != 1:
        raise ValueError(f"{name} must be a scalar or have shape [B]")
    return a


def rms_norm(x):
    """optimistix.rms_norm: sqrt(mean(x^2)) (the only norm the kernels implement)."""
    import torch
    x = torch.as_tensor(x)
    if x.numel() == 1:
        return x.abs().reshape(())
    return torch.sqrt((x * x).sum()) / math.sqrt(x.numel())


def adjoint_rms_seminorm(x):
    """The adjoint seminorm (_misc.py:85-116): the RMS norm of (y, a_y) of a backward
    state (y, a_y, a_args, a_terms), leaving the parameter adjoints out of the error
    control.  For PIDController(norm=adjoint_rms_seminorm) as the stepsize_controller of
    BacksolveAdjoint; the neural-ODE backsolve kernel implements it."""
    assert isinstance(x, tuple) and len(x) == 4
    import torch
    y, a_y = (torch.as_tensor(v).reshape(-1) for v in x[:2])
    return rms_norm(torch.cat([y, a_y]))


class AbstractStepSizeController:
    def wrap(self, direction):
        return self


class AbstractAdaptiveStepSizeController(AbstractStepSizeController):
    pass


class ConstantStepSize(AbstractStepSizeController):
    """Constant step size dt0 (constant.py:20-104): steps land on
    t0 + (t1 - t0) * k / N with N = ceil((t1 - t0) / nextafter(dt0))."""

    def __repr__(self):
        return "ConstantStepSize()"


class _MetaPID(type):
    """PIDController(..., step_ts=..., jump_ts=...) builds the ClipStepSizeController
    around a plain PIDController, as pid.py:91-97 does."""

    def __call__(cls, *args, **kwargs):
        step_ts = kwargs.pop("step_ts", None)
        jump_ts = kwargs.pop("jump_ts", None)
        if step_ts is not None or jump_ts is not None:
            return ClipStepSizeController(cls(*args, **kwargs), step_ts, jump_ts)
        return super().__call__(*args, **kwargs)


class PIDController(AbstractAdaptiveStepSizeController, metaclass=_MetaPID):
    """PID step-size controller with the fields and defaults of pid.py:299-311."""

    def __init__(self, rtol, atol, norm: Callable = rms_norm, pcoeff=0, icoeff=1, dcoeff=0,
                 dtmin: Optional[float] = None, dtmax: Optional[float] = None,
                 force_dtmin: bool = True, factormin=0.2, factormax=10.0, safety=0.9,
                 error_order: Optional[float] = None):
        if rtol is None or atol is None:
            raise ValueError(
                "The default values for `rtol` and `atol` were removed in Diffrax "
                "version 0.1.0. (As the choice of tolerance is nearly always "
                "something that you, as an end user, should make an explicit choice "
                "about.)")
        if norm is not rms_norm and norm is not adjoint_rms_seminorm:
            raise NotImplementedError(
                "diffrax_amd kernels implement rms_norm and adjoint_rms_seminorm only")
        self.rtol = _tolerance(rtol, "rtol")
        self.atol = _tolerance(atol, "atol")
        self.norm = norm
        self.pcoeff = pcoeff
        self.icoeff = icoeff
        self.dcoeff = dcoeff
        self.dtmin = dtmin
        self.dtmax = dtmax
        self.force_dtmin = bool(force_dtmin)
        self.factormin = float(factormin)
        self.factormax = float(factormax)
        self.safety = float(safety)
        self.error_order = error_order

    def __repr__(self):
        return f"PIDController(rtol={self.rtol}, atol={self.atol})"


def _none_or_sorted(x):
    if x is None:
        return None
    import numpy as np
    return np.sort(np.asarray(x.detach().cpu() if hasattr(x, "detach") else x, dtype=np.float64))


class ClipStepSizeController(AbstractAdaptiveStepSizeController):
    """Clips the steps of an adaptive controller so that they land exactly on `step_ts`
    and stop just before / restart just after `jump_ts` (clip.py:120-428).  The wrapped
    controller must be a PIDController; `store_rejected_steps` (revisiting rejected steps
    for SDEs with non-commutative noise) is not provided."""

    def __init__(self, controller, step_ts=None, jump_ts=None, store_rejected_steps=None,
                 _callback_on_reject=None):
        if not isinstance(controller, AbstractAdaptiveStepSizeController):
            raise ValueError(
                "Can only apply `ClipStepSizeController` to adaptive step size "
                f"controllers, but got {controller}.")
        if (store_rejected_steps is not None) and (store_rejected_steps <= 0):
            raise ValueError(
                "`store_rejected_steps must either be `None` or a non-negative integer.")
        self.controller = controller
        self.step_ts = _none_or_sorted(step_ts)
        self.jump_ts = _none_or_sorted(jump_ts)
        self.store_rejected_steps = store_rejected_steps
        self.callback_on_reject = _callback_on_reject

    @property
    def rtol(self):
        return self.controller.rtol

    @property
    def atol(self):
        return self.controller.atol

    @property
    def norm(self):
        return self.controller.norm

    def __repr__(self):
        return (f"ClipStepSizeController({self.controller!r}, step_ts={self.step_ts}, "
                f"jump_ts={self.jump_ts})")
