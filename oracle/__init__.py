"""FP64 CPU oracle for FourierFlows.jl's pseudospectral step.

TEST INFRASTRUCTURE ONLY.  This package restates, in NumPy/SciPy (FP64,
pocketfft), the arithmetic of the reference's hot path so that the HIP
library can be checked against it.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or the timed CPU baseline -- never as part of the product path.

Modules (each function cites the reference file:line it follows):

* ``grids``        -- OneDGrid/TwoDGrid/ThreeDGrid, wavenumbers, dealias,
                      ``mul!``/``ldiv!`` conventions        (src/domains.jl)
* ``timesteppers`` -- getexpLs/getetdcoeffs, ETDRK4, RK4, Problem, Clock
                                                     (src/timesteppers.jl,
                                                      src/problem.jl)
* ``physics``      -- 2D vorticity and 3D tracer nonlinear terms, the 1D
                      diffusion testbed                  (src/diffusion.jl,
                      pattern src/diffusion.jl:135-143, formula
                      src/utils.jl:190-205)
* ``utils``        -- parsevalsum/parsevalsum2, jacobian (src/utils.jl)
* ``ics``          -- SplitMix64 and the seeded initial conditions
* ``adjoint``      -- discrete adjoint of the ETDRK4/RK4 steps (no
                      counterpart in the reference)

Every function is pinned by checks in ``tests/test_oracle*.py`` that do not
restate its formula: the reference's known-answer tests, brute-force DFTs,
mpmath, closed forms, convergence orders and dot-product identities.
"""

from . import grids, timesteppers, physics, utils, ics  # noqa: F401
