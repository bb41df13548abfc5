"""diffrax_amd -- diffrax's explicit Runge-Kutta path as MI355X (gfx950) HIP kernels.

Drop-in objects with the reference names and defaults (diffrax 0.7.2):
diffeqsolve, ODETerm, ControlTerm, MultiTerm, Tsit5, Dopri5, Heun, PIDController,
ConstantStepSize, SaveAt, SubSaveAt, RecursiveCheckpointAdjoint, BacksolveAdjoint, RESULTS,
Solution.
Vector fields are registered device functions (LotkaVolterra, VanDerPol, LinearField,
MLPField, OUDrift/DiagonalNoise); the integration loop runs in persistent HIP kernels
behind the C ABI of include/diffrax_amd.h.  There is no CPU fallback.
"""
from ._adjoint import AbstractAdjoint, BacksolveAdjoint, RecursiveCheckpointAdjoint
from ._event import AffineCondition, Event, Newton
from ._integrate import ErkPlan, diffeqsolve
from ._saveat import SaveAt, SubSaveAt, save_y
from ._solution import RESULTS, Solution, is_okay, is_successful
from ._solver import (AbstractAdaptiveSolver, AbstractERK, AbstractSolver, ButcherTableau,
                      Dopri5, Heun, Tsit5)
from ._step_size_controller import (AbstractAdaptiveStepSizeController,
                                    AbstractStepSizeController, ClipStepSizeController,
                                    ConstantStepSize, PIDController, adjoint_rms_seminorm,
                                    rms_norm)
from ._term import (AbstractPath, AbstractTerm, ControlTerm, IncrementPath, MultiTerm,
                    ODETerm)
from ._vector_fields import (DiagonalNoise, LinearField, LotkaVolterra, MLPField, OUDrift,
                             RegisteredVectorField, VanDerPol)
from . import _mlp  # noqa: F401  (registers the neural-ODE C-ABI signatures)
from . import _sde  # noqa: F401  (registers the SDE C-ABI signatures)

__version__ = "0.1.0"
