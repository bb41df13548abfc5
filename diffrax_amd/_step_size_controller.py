"""Step-size controllers (mirrors diffrax/_step_size_controller/{base,pid,constant}.py).

These are configuration objects: the accept/reject logic itself runs inside the device
kernels (erk_common.hpp), one lane per trajectory."""
from __future__ import annotations

import math
from typing import Callable, Optional


def rms_norm(x):
    """optimistix.rms_norm: sqrt(mean(x^2)) (the only norm the kernels implement)."""
    import torch
    x = torch.as_tensor(x)
    if x.numel() == 1:
        return x.abs().reshape(())
    return torch.sqrt((x * x).sum()) / math.sqrt(x.numel())


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


class PIDController(AbstractAdaptiveStepSizeController):
    """PID step-size controller with the fields and defaults of pid.py:299-311."""

    def __init__(self, rtol, atol, norm: Callable = rms_norm, pcoeff=0, icoeff=1, dcoeff=0,
                 dtmin: Optional[float] = None, dtmax: Optional[float] = None,
                 force_dtmin: bool = True, factormin=0.2, factormax=10.0, safety=0.9,
                 error_order: Optional[float] = None, step_ts=None, jump_ts=None):
        if rtol is None or atol is None:
            raise ValueError(
                "The default values for `rtol` and `atol` were removed in Diffrax "
                "version 0.1.0. (As the choice of tolerance is nearly always "
                "something that you, as an end user, should make an explicit choice "
                "about.)")
        if step_ts is not None or jump_ts is not None:
            raise NotImplementedError(
                "step_ts/jump_ts (ClipStepSizeController) are not provided by diffrax_amd")
        if norm is not rms_norm:
            raise NotImplementedError("diffrax_amd kernels implement rms_norm only")
        self.rtol = float(rtol)
        self.atol = float(atol)
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
