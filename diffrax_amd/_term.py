"""Terms and paths (mirrors diffrax/_term.py:35-764 and diffrax/_path.py).

A term's vector field must be one of the registered device vector fields
(diffrax_amd._vector_fields); the kernels evaluate it on the GPU.  The Python
``vf/contr/prod`` methods evaluate the same formulas with torch for inspection."""
from __future__ import annotations

from typing import Any


class AbstractTerm:
    def vf(self, t, y, args):
        raise NotImplementedError

    def contr(self, t0, t1, **kwargs):
        raise NotImplementedError

    def prod(self, vf, control):
        raise NotImplementedError

    def vf_prod(self, t, y, args, control):
        return self.prod(self.vf(t, y, args), control)

    def is_vf_expensive(self, t0, t1, y, args) -> bool:
        return False


class ODETerm(AbstractTerm):
    """f(t, y, args) dt (_term.py:174-228)."""

    def __init__(self, vector_field):
        self.vector_field = vector_field

    def vf(self, t, y, args):
        return self.vector_field(t, y, args)

    def contr(self, t0, t1, **kwargs):
        return t1 - t0

    def prod(self, vf, control):
        return control * vf


class AbstractPath:
    @property
    def t0(self):
        raise NotImplementedError

    @property
    def t1(self):
        raise NotImplementedError

    def evaluate(self, t0, t1=None, left: bool = True):
        raise NotImplementedError


class IncrementPath(AbstractPath):
    """A control path given by host-generated increments on a fixed step grid.

    ``dW`` has shape [num_steps, batch, d] (or [num_steps, d] for one path); step n of a
    ConstantStepSize solve uses ``dW[n]`` as its control increment.  This is how the
    reference feeds externally generated Brownian increments through a callable
    (t0, t1) -> dW (_term.py:238-264)."""

    def __init__(self, dW, t0: float = 0.0, t1: float = float("inf")):
        self.dW = dW
        self._t0 = t0
        self._t1 = t1

    @property
    def t0(self):
        return self._t0

    @property
    def t1(self):
        return self._t1


class ControlTerm(AbstractTerm):
    """g(t, y, args) dX (_term.py:271-576); with a diagonal g the product is g * dX."""

    def __init__(self, vector_field, control: Any):
        self.vector_field = vector_field
        self.control = control

    def vf(self, t, y, args):
        return self.vector_field(t, y, args)

    def contr(self, t0, t1, **kwargs):
        return self.control.evaluate(t0, t1, **kwargs)

    def prod(self, vf, control):
        return vf * control


class MultiTerm(AbstractTerm):
    """Sum of terms (_term.py:666-731): vf_prod = sum_i term_i.vf_prod."""

    def __init__(self, *terms: AbstractTerm):
        self.terms = tuple(terms)

    def vf(self, t, y, args):
        return tuple(term.vf(t, y, args) for term in self.terms)

    def contr(self, t0, t1, **kwargs):
        return tuple(term.contr(t0, t1, **kwargs) for term in self.terms)

    def prod(self, vf, control):
        out = [term.prod(v, c) for term, v, c in zip(self.terms, vf, control)]
        total = out[0]
        for o in out[1:]:
            total = total + o
        return total


def is_sde(terms) -> bool:
    if isinstance(terms, MultiTerm):
        return any(is_sde(t) for t in terms.terms)
    return isinstance(terms, ControlTerm) and isinstance(terms.control, IncrementPath)
