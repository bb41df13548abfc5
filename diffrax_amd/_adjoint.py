"""Adjoint selection (mirrors diffrax/_adjoint.py:118-329 and 539-873).

RecursiveCheckpointAdjoint is the discretise-then-optimise gradient: reverse mode through
the accepted step sequence, with the step-size factors treated as constants
(pid.py:498-530).  In diffrax_amd it is realised by hand-written backward kernels that
replay the recorded accepted-step grid from sqrt(N)-spaced checkpoints.

BacksolveAdjoint is the optimise-then-discretise gradient: the continuous adjoint
equations solved backwards in time between consecutive save times, one lane per
instance, by dfx_erk_backsolve."""
from __future__ import annotations

from typing import Optional


class AbstractAdjoint:
    pass


class RecursiveCheckpointAdjoint(AbstractAdjoint):
    """Backpropagate through the solver with checkpointed recomputation.

    checkpoints: number of stored states per trajectory.  The small-system solves
    (LotkaVolterra, VanDerPol, LinearField) store the full step state every
    K = ceil(max_steps / checkpoints) accepted steps, checkpoints = ceil(sqrt(max_steps))
    when None; the backward pass replays each K-step segment from its checkpoint and
    sweeps it in reverse, so memory is O(checkpoints + K) per trajectory.  The neural-ODE
    solve stores a state every K accepted steps by the same rule.  The reference picks ~log(max_steps) checkpoints for its
    online binomial scheme (_adjoint.py:315-329); both bound memory, the device kernels
    use uniform spacing."""

    def __init__(self, checkpoints: Optional[int] = None):
        self.checkpoints = checkpoints

    def __repr__(self):
        return f"RecursiveCheckpointAdjoint(checkpoints={self.checkpoints})"


class BacksolveAdjoint(AbstractAdjoint):
    """Gradients by the continuous adjoint (BacksolveAdjoint, _adjoint.py:738-873).

    From the last saved time back to t0 the augmented state (y, a_y, a_args) follows
    (f, -a_y^T df/dy, -a_y^T df/dargs) (AdjointTerm, _term.py:767-917), solved between
    consecutive save times with y restarting from the saved state and a_y receiving each
    save's cotangent (_adjoint.py:646-735).  Gradients flow to y0 and to the vector
    field's args; t0, t1 and the save times get none (the reference raises for them).

    kwargs: the options of the backward solves, as in the reference: ``dt0`` (used as
    given; the default is the forward's -dt0), ``solver`` (Tsit5 / Dopri5 / Heun),
    ``stepsize_controller`` (PIDController with integral control and no dtmin, or
    ConstantStepSize), ``max_steps`` (per interval) and ``throw``; ``adjoint`` is accepted
    and unused (the backward solves are not differentiated again)."""

    _valid = {"dt0", "solver", "stepsize_controller", "adjoint", "max_steps", "throw"}

    def __init__(self, **kwargs):
        bad = set(kwargs) - self._valid
        if bad:
            raise ValueError("The following keyword arguments are not valid for "
                             f"`BacksolveAdjoint`: {bad}")
        self.kwargs = kwargs

    def __repr__(self):
        return f"BacksolveAdjoint({', '.join(f'{k}={v!r}' for k, v in self.kwargs.items())})"
