"""Solvers (mirrors diffrax/_solver/{base,runge_kutta,tsit5,dopri5,heun}.py).

The solver objects select the device tableau; their coefficients live in the HIP
library (erk_common.hpp) and are read back through dfx_tableau for inspection."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass(frozen=True)
class ButcherTableau:
    """Explicit Butcher tableau (runge_kutta.py:48-166): c excludes c1 = 0."""

    c: np.ndarray
    b_sol: np.ndarray
    b_error: np.ndarray
    a_lower: tuple
    fsal: bool
    ssal: bool
    num_stages: int


def _device_tableau(solver_id: int) -> ButcherTableau:
    lib = _lib.load()
    a = (ctypes.c_double * 49)()
    b = (ctypes.c_double * 7)()
    be = (ctypes.c_double * 7)()
    c = (ctypes.c_double * 7)()
    info = (ctypes.c_int32 * 4)()
    _lib.check(lib.dfx_tableau(solver_id, a, b, be, c, info), "dfx_tableau")
    S = info[0]
    A = np.array(a[:]).reshape(7, 7)
    return ButcherTableau(
        c=np.array(c[1:S]), b_sol=np.array(b[:S]), b_error=np.array(be[:S]),
        a_lower=tuple(A[i, :i].copy() for i in range(1, S)), fsal=bool(info[1]),
        ssal=bool(info[2]), num_stages=S)


class AbstractSolver:
    solver_id: int = -1

    def order(self, terms=None):
        raise NotImplementedError

    def strong_order(self, terms=None):
        return None

    def error_order(self, terms=None):
        """Error order for adaptive stepping (_solver/base.py:97-121)."""
        from ._term import is_sde
        if terms is not None and is_sde(terms):
            order = self.strong_order(terms)
            return None if order is None else order + 0.5
        return self.order(terms)

    @property
    def tableau(self) -> ButcherTableau:
        return _device_tableau(self.solver_id)

    def __repr__(self):
        return f"{type(self).__name__}()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))


class AbstractAdaptiveSolver(AbstractSolver):
    pass


class AbstractERK(AbstractAdaptiveSolver):
    """Explicit Runge-Kutta family handled by the device kernels."""


class Tsit5(AbstractERK):
    """Tsitouras' 5/4 method: 7 stages, FSAL, 5th order dense output."""

    solver_id = _lib.TSIT5

    def order(self, terms=None):
        return 5


class Dopri5(AbstractERK):
    """Dormand-Prince 5/4 method (Shampine's tableau): 7 stages, FSAL, quartic dense
    output."""

    solver_id = _lib.DOPRI5

    def order(self, terms=None):
        return 5


class Heun(AbstractERK):
    """Heun's method: 2 stages, embedded Euler, cubic Hermite dense output;
    Stratonovich for SDEs."""

    solver_id = _lib.HEUN

    def order(self, terms=None):
        return 2

    def strong_order(self, terms=None):
        return 0.5
