"""fourierflows_hip -- FourierFlows.jl's pseudospectral step on AMD MI355X.

A Python host layer over libffhip.so (HIP/C++ for gfx950, C ABI in
include/ffhip.h) that keeps the reference's API names:

* grids: ``OneDGrid``, ``TwoDGrid``, ``ThreeDGrid`` with ``rfftplan`` /
  ``fftplan``, ``mul`` (``mul!``), ``ldiv`` (``ldiv!``), ``dealias``
  (``dealias!``), ``getaliasedwavenumbers``        (src/domains.jl)
* ``Equation``, ``Clock``, ``Problem``, ``TimeStepper``, the ETDRK4, RK4,
  ForwardEuler, AB3 and LSRK54 steppers and their Filtered variants
  (``makefilter``), ``stepforward`` (``stepforward!``), ``step_until``
                                           (src/problem.jl, src/timesteppers.jl)
* ``parsevalsum2``, ``parsevalsum``               (src/utils.jl)
* ``Diagnostic``, ``increment``, ``update`` and ``stepforward(prob, diags,
  nsteps)``                                        (src/diagnostics.jl)
* ``Output``, ``saveproblem``, ``saveoutput``, ``savediagnostic`` (npz
  archives with the reference's group layout)     (src/output.jl)
* built-in operators ``VorticityOperator`` / ``TracerOperator`` and the
  ``Diffusion`` testbed                              (src/diffusion.jl)
* ``adjoint``: differentiable ``ldiv``/``mul``/``dealias`` and the discrete
  ETDRK4 adjoint (no counterpart in the reference)

Device arrays are torch tensors on the ROCm device (``HIP()`` is the device
tag, the analogue of ``GPU()``).
"""

from . import _lib, adjoint, diagnostics, output
from .diagnostics import Diagnostic, increment, update
from .output import Output, savediagnostic, saveoutput, saveproblem
from ._lib import DomainError, FFHIPError, NonFiniteError
from .core import CPU, GPU, HIP, cxtype, device_array, fltype, init, zeros
from .domains import (OneDGrid, Plan, ThreeDGrid, TwoDGrid, dealias, getaliasedwavenumbers, gridpoints, ldiv,
                      makefilter, mul)
from .ics import ensemble_members, splitmix64, tracer_ic, uniform01, vorticity_ic
from .physics import (Diffusion, TracerOperator, TracerProblem, TwoDVorticityProblem, VorticityEnsemble,
                      VorticityOperator, vorticity_linear)
from .timesteppers import (AB3TimeStepper, Clock, EmptyParams, EmptyVars, Equation, ETDRK4TimeStepper,
                           FilteredAB3TimeStepper, FilteredETDRK4TimeStepper, FilteredForwardEulerTimeStepper,
                           FilteredLSRK54TimeStepper, FilteredRK4TimeStepper, ForwardEulerTimeStepper,
                           LSRK54TimeStepper, Problem, RK4TimeStepper, TimeStepper, isexplicit, step_until,
                           stepforward)
from .utils import etdcoeffs, parsevalsum, parsevalsum2

# Julia-style aliases
stepforward_b = stepforward
dealias_b = dealias

__all__ = [n for n in dir() if not n.startswith("_")]


def load_library():
    """Load libffhip.so (raises if it has not been built)."""
    return _lib.load()
