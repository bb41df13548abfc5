"""Adjoint selection (mirrors diffrax/_adjoint.py:118-329).

RecursiveCheckpointAdjoint is the discretise-then-optimise gradient: reverse mode through
the accepted step sequence, with the step-size factors treated as constants
(pid.py:498-530).  In diffrax_amd it is realised by hand-written backward kernels that
replay the recorded accepted-step grid from sqrt(N)-spaced checkpoints."""
from __future__ import annotations

from typing import Optional


class AbstractAdjoint:
    pass


class RecursiveCheckpointAdjoint(AbstractAdjoint):
    """Backpropagate through the solver with checkpointed recomputation.

    checkpoints: number of stored states per trajectory; None picks ceil(sqrt(N)) for N
    accepted steps (the reference picks ~log(max_steps) for its online binomial scheme,
    _adjoint.py:315-329; both bound memory, the device kernels use uniform sqrt(N))."""

    def __init__(self, checkpoints: Optional[int] = None):
        self.checkpoints = checkpoints

    def __repr__(self):
        return f"RecursiveCheckpointAdjoint(checkpoints={self.checkpoints})"
